"""Temporal blocking of the collide pass (lg_set_temporal_depth) against the
oracle, through the C ABI (gpu): tiles of rows stepped several steps in
registers, their edge rows recomputed, give the CPU path's cells bit for bit
at every depth, on ragged shapes (the ds_bpermute path for rows under 64
words), on 64-word rows (the DPP path), on the reference's golden cases and
on the 9.4M-cell plate workload. Each case prints what it is first."""
import sys

import numpy as np
import pytest

from lgca_amd import GpuLattice
import workloads as W
from oracle.workloads_cpu import config_oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(oracle, w, h, rest, bands, inlet, seed, depth, rows, ahead, calls, density=0.22,
        inflow=(0.3, 0.0), disc=True):
    say(f"temporal {w}x{h} rest={rest} bands={bands} inlet={inlet} depth={depth} rows={rows} "
        f"ahead={ahead} calls={calls}")
    lat = oracle.Lattice(w, h, density, inflow, inlet, rest, bands, seed)
    if disc:
        yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
        lat.add_solid_disc(w / 3.0, yc, 3.5)
    lat.init_equilibrium()
    g = GpuLattice(w, h, rest, seed, bands)
    g.set_inlet(inlet, density, inflow)
    g.load(lat.cells)
    g.set_draw_ahead(ahead)
    g.set_temporal_depth(depth, rows)
    for n in calls:
        lat.advance(n)
        g.advance(n)
        got = g.store()
        if not np.array_equal(got, lat.cells):
            bad = np.nonzero(got != lat.cells)[0]
            pytest.fail(f"after {lat.steps} steps: {bad.size} cells differ, first at "
                        f"{[(int(b) % w, int(b) // w) for b in bad[:4]]}")
    assert g.steps == lat.steps
    g.close()


SHAPES = [
    (65, 6, True, 1, 8),        # tiles taller than the lattice
    (100, 18, True, 3, 8),
    (129, 64, False, 4, 70),    # inflow wider than a word, no rest
    (200, 40, True, 7, 130),
    (64, 32, True, 5, 0),       # one word per row
    (1000, 26, True, 2, 8),     # 16 words per row
]


@pytest.mark.parametrize("w,h,rest,bands,inlet", SHAPES)
@pytest.mark.parametrize("depth,rows,ahead", [(2, 0, 2), (3, 0, 7), (4, 3, 8), (8, 0, 8),
                                              (8, 16, 16)])
def test_ragged_shapes(oracle, w, h, rest, bands, inlet, depth, rows, ahead):
    run(oracle, w, h, rest, bands, inlet, 0xB10C + w, depth, rows, ahead, (1, 9, 16, 5))


@pytest.mark.parametrize("w", [4096, 4033, 4001])
@pytest.mark.parametrize("depth,rows", [(2, 0), (4, 0), (8, 0), (4, 24)])
def test_full_wave_rows(oracle, w, depth, rows):
    """64 words per row: neighbours across lanes by whole-wave DPP rotation
    (a ragged last word at 4033 and 4001 cells)."""
    run(oracle, w, 48, True, 3, 8, 0xF00 + w, depth, rows, 2 * depth, (2 * depth, 3, 8))


@pytest.mark.parametrize("case", [c[0] for c in W.GOLDEN_CASES])
@pytest.mark.parametrize("depth", [2, 4, 8])
def test_golden_checksums(oracle, case, depth):
    """tests/golden.rs:141-147 with temporal blocking: the reference's cell
    checksum after 200 steps."""
    c = next(c for c in W.GOLDEN_CASES if c[0] == case)
    name, w, h, rest, inlet, plate, threads, steps = c
    say(f"golden {name} depth={depth}")
    lat = oracle.golden_build(c)
    lat.advance(steps)
    g = W.golden_gpu(c)
    g.set_draw_ahead(depth)
    g.set_temporal_depth(depth)
    g.advance(steps)
    assert np.array_equal(g.store(), lat.cells)
    g.close()


@pytest.mark.parametrize("depth", [2, 4, 8])
def test_plate_workload_9m_cells(oracle, depth):
    """The 4096 x 2304 plate workload (16 bands, 8 inflow columns) for 60
    steps in calls of 20, against the oracle's replay."""
    cfg = W.config2(bands=16)
    say(f"config2 depth={depth}")
    lat = config_oracle(oracle, cfg)
    lat.set_parallel(True)
    g = W.config_gpu(cfg)
    g.set_draw_ahead(depth)
    g.set_temporal_depth(depth)
    for _ in range(3):
        lat.advance(20)
        g.advance(20)
        assert np.array_equal(g.store(), lat.cells)
    g.close()


@pytest.mark.parametrize("w,h,bands,inlet", [(4096, 2304, 16, 8), (320, 600, 16, 8), (1000, 26, 2, 70)])
def test_autotune_keeps_results_exact(oracle, w, h, bands, inlet):
    """lg_autotune tries every pipeline this lattice can run (resident bands,
    one-launch two-pass, draw-ahead, fused, temporally blocked collides of
    depth 2, 4 and 8) on the lattice itself; its steps and the steps after
    it, in the pipeline it keeps, equal the CPU path's."""
    seed = 0xA70 + w
    print(f"autotune {w}x{h} bands={bands} inlet={inlet}", flush=True)
    lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), inlet, True, bands, seed)
    yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
    lat.add_solid_disc(w / 3.0, yc, 3.5)
    lat.init_equilibrium()
    lat.set_parallel(True)
    g = GpuLattice(w, h, True, seed, bands)
    g.set_inlet(inlet, 0.22, (0.3, 0.0))
    g.load(lat.cells)
    used = g.autotune(0)
    k, _ = g.temporal_depth()
    print(f"  used {used} steps, pipeline {g.pipeline()}, resident {g.resident()}, depth {k}",
          flush=True)
    assert used > 0 and used % 32 == 0
    g.advance(17)
    lat.advance(used + 17)
    assert g.steps == lat.steps
    assert np.array_equal(g.store(), lat.cells)
    g.close()
