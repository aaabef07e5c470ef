"""Frames straight from the engine (gpu): the run loop of src/main.rs:669-709
on a lattice, each frame queued by lg_frames_write_lattice, against the
oracle's frames of its own CPU replay of the same run."""
import sys

import numpy as np
import pytest

import lgca_amd as L
import workloads as W
from oracle.workloads_cpu import config_oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("workers", [1, 4])
def test_run_loop_frames_match_the_cpu_replay(oracle, tmp_path, workers):
    """512x256 plate run: three frames 100 steps apart, a field sample every
    5 steps at alpha 0.2, blocks of 16, scale 8, arrows 1.6 in the frame of
    the field's mean velocity. Every file equals the oracle's."""
    cfg = W.config1()
    every, alpha, frame_every, bx = 5, 0.2, 100, 16
    say(f"frames config1 workers={workers}")
    lat = config_oracle(oracle, cfg)
    g = W.config_gpu(cfg)
    g.attach_field(bx, bx)
    assert g.field_dims() == (cfg["w"] // bx, cfg["h"] // bx, bx, bx)
    f = oracle.Field(cfg["w"], cfg["h"], bx, bx)
    bw, bh = f.bw, f.bh
    want = []
    with L.Frames(workers) as fr:
        for k in range(3):
            g.advance_sampling(frame_every, every, alpha)
            fr.write_lattice(g, tmp_path / f"vorticity-{k:04}.png", tmp_path / f"arrows-{k:04}.svg", 8, 1.6)
            for _ in range(frame_every // every):
                lat.advance(every)
                f.sample(lat, alpha)
            ux, uy, rho, solid = f.read()
            vort = oracle.vorticity(ux, uy, bw, bh, bx, bx)
            wmax = oracle.vorticity_scale(vort, solid)
            mean = oracle.field_mean_velocity(ux, uy, solid)
            want.append((oracle.vorticity_png(ux, uy, solid, bw, bh, bx, bx, 8, wmax),
                         oracle.arrows_svg(ux, uy, solid, bw, bh, bx, bx, 1.6, mean)))
    for k, (png, svg) in enumerate(want):
        assert (tmp_path / f"vorticity-{k:04}.png").read_bytes() == png
        assert (tmp_path / f"arrows-{k:04}.svg").read_bytes() == svg


def test_write_lattice_needs_a_field(tmp_path):
    say("frames without a field")
    g = L.GpuLattice(64, 32, True, 1, 1)
    with L.Frames(1) as fr:
        with pytest.raises(L.LgcaError, match="no field"):
            fr.write_lattice(g, tmp_path / "v.png", tmp_path / "a.svg", 8, 1.6)
