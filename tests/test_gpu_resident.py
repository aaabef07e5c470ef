"""Resident bands (lg_set_resident) against the oracle, through the C ABI
(gpu): a persistent cooperative launch per chunk in which each band of rows
stays in registers and meets its neighbours through halo rows and step
counters gives the CPU path's cells bit for bit -- rows under 64 words (the
ds_bpermute path) and of 64 words (DPP), one-row bands, a short last band,
calls of every length (chunks cut at 256 steps and at samples), the
reference's golden checksums and the 9.4M-cell plate workload. Each case
prints what it is first."""
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


def run(oracle, w, h, rest, bands, inlet, seed, calls, density=0.22, inflow=(0.3, 0.0), every=0):
    say(f"resident {w}x{h} rest={rest} bands={bands} inlet={inlet} calls={calls} every={every}")
    lat = oracle.Lattice(w, h, density, inflow, inlet, rest, bands, seed)
    yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
    lat.add_solid_disc(w / 3.0, yc, 3.5)
    lat.init_equilibrium()
    g = GpuLattice(w, h, rest, seed, bands)
    g.set_inlet(inlet, density, inflow)
    g.load(lat.cells)
    g.set_resident(True)
    assert g.resident() is not None
    for n in calls:
        lat.advance(n)
        if every:
            g.advance_sampling(n, every, 1.0)
        else:
            g.advance(n)
        got = g.store()
        if not np.array_equal(got, lat.cells):
            bad = np.nonzero(got != lat.cells)[0]
            pytest.fail(f"after {lat.steps} steps: {bad.size} cells differ, first at "
                        f"{[(int(b) % w, int(b) // w) for b in bad[:4]]}")
    assert g.steps == lat.steps
    g.close()


@pytest.mark.parametrize("w,h,rest,bands,inlet", [
    (65, 6, True, 1, 8),        # one-row bands, tiny height
    (100, 18, True, 3, 8),
    (129, 64, False, 4, 70),    # inflow wider than a word, no rest
    (200, 40, True, 7, 130),
    (64, 32, True, 5, 0),       # one word per row
    (1000, 26, True, 2, 8),
    (256, 2306, True, 16, 8),   # ten-row bands, a six-row last band
    (320, 600, True, 16, 8),    # three-row bands
    (192, 1200, True, 16, 8),   # five-row bands: the first row's draws on two helper waves
    (130, 2250, False, 5, 70),  # nine-row bands with helpers, no rest, wide inflow
])
def test_ragged_shapes(oracle, w, h, rest, bands, inlet):
    run(oracle, w, h, rest, bands, inlet, 0xBA5E + w, (1, 2, 9, 16, 5))


@pytest.mark.parametrize("w", [4096, 4033, 4001])
def test_full_wave_rows(oracle, w):
    """64 words per row: neighbours across lanes by whole-wave DPP rotation."""
    run(oracle, w, 96, True, 3, 8, 0xF00 + w, (3, 8, 13))


def test_chunks_across_tables_and_samples(oracle):
    """Calls longer than the 256-step table chunk, and chunks cut by samples."""
    run(oracle, 192, 64, True, 4, 8, 0x51, (300, 40), every=0)
    run(oracle, 192, 64, True, 4, 8, 0x52, (130, 45), every=50)


@pytest.mark.parametrize("case", [c[0] for c in W.GOLDEN_CASES])
def test_golden_checksums(oracle, case):
    """tests/golden.rs:141-147 with resident bands: the reference's cell
    checksum after 200 steps."""
    c = next(c for c in W.GOLDEN_CASES if c[0] == case)
    name, w, h, rest, inlet, plate, threads, steps = c
    say(f"golden {name} resident")
    lat = oracle.golden_build(c)
    lat.advance(steps)
    g = W.golden_gpu(c)
    g.set_resident(True)
    g.advance(steps)
    assert np.array_equal(g.store(), lat.cells)
    g.close()


def test_plate_workload_9m_cells(oracle):
    """The 4096 x 2304 plate workload (16 bands, 8 inflow columns) for 60
    steps in calls of 20 with the 32 x 32 field sampled every 100 steps,
    against the oracle's replay (cells and exact block sums)."""
    cfg = W.config2(bands=16)
    say("config2 resident")
    lat = config_oracle(oracle, cfg)
    lat.set_parallel(True)
    g = W.config_gpu(cfg)
    g.set_resident(True)
    assert g.resident() == (9, 256)
    g.attach_field(32, 32)
    f = oracle.Field(cfg["w"], cfg["h"], 32, 32)
    for _ in range(3):
        lat.advance(20)
        f.sample(lat, 1.0)
        g.advance_sampling(20, 100, 1.0)
        assert np.array_equal(g.store(), lat.cells)
    for a, b in zip(g.read_block_sums(), f.block_sums(lat)):
        assert np.array_equal(a, b)
    for a, b in zip(g.read_field(), f.read()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    g.close()


def replay_sampling(lat, f, n, every, alpha):
    """What one advance_sampling(n, every, alpha) call does (src/sim.rs:119-143):
    a sample after every `every` steps of the call and one at its end."""
    left, since = n, 0
    while left:
        c = min(left, every - since)
        lat.advance(c)
        left -= c
        since += c
        if since >= every or left == 0:
            since = 0
            f.sample(lat, alpha)


@pytest.mark.parametrize("w,h,bands,bx,by", [
    (256, 2306, 16, 20, 20),   # ten-row bands, two bands per block row, blocks across words
    (320, 600, 16, 7, 3),      # three-row bands, as tall as a block row
    (1000, 26, 2, 100, 5),     # one-row bands, blocks wider than a word
    (4096, 96, 3, 64, 40),     # 64-word rows, forty bands per block row, a partial block row
    (129, 700, 4, 43, 9),      # a partial block column
    (4096, 96, 3, 1, 1),       # blocks too many for a band's LDS: the separate kernel
])
def test_coarse_graining_in_the_band_launch(oracle, w, h, bands, bx, by):
    """A chunk that ends at a sample coarse-grains its final state in the
    band launch itself: block sums and the f32 running average
    (src/render.rs:93-115, alpha 0.3) bit for bit with the oracle's, over
    calls that sample inside and at their end."""
    seed = 0xF1E1D + w + bx
    say(f"fused sample {w}x{h} bands={bands} blocks {bx}x{by}")
    lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, bands, seed)
    yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
    lat.add_solid_disc(w / 3.0, yc, 5.5)
    lat.init_equilibrium()
    g = GpuLattice(w, h, True, seed, bands)
    g.set_inlet(8, 0.22, (0.3, 0.0))
    g.load(lat.cells)
    g.set_resident(True)
    assert g.resident() is not None
    g.attach_field(bx, by)
    f = oracle.Field(w, h, bx, by)
    for n, every in ((20, 100), (25, 10), (7, 3), (40, 40)):
        replay_sampling(lat, f, n, every, 0.3)
        g.advance_sampling(n, every, 0.3)
        assert np.array_equal(g.store(), lat.cells), f"cells after {lat.steps} steps"
        for name, a, b in zip(("ux", "uy", "rho", "solid"), g.read_field(), f.read()):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
                f"field {name} after {lat.steps} steps"
        for a, b in zip(g.read_block_sums(), f.block_sums(lat)):
            assert np.array_equal(a, b), f"block sums after {lat.steps} steps"
    g.close()


def test_concurrent_lattices_on_their_own_streams(oracle):
    """Four lattices whose advances are queued at once on their own streams,
    each with resident-band launches of its own: every one equals its CPU
    replay (the runtime may run the cooperative launches one after another,
    or decline one, in which case that lattice takes another pipeline)."""
    say("concurrent resident lattices")
    shapes = [(4096, 2304, 16), (2048, 1280, 8), (256, 256, 4), (1000, 26, 2)]
    runs = []
    for w, h, bands in shapes:
        seed = 0xC0C0 + w
        lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, bands, seed)
        lat.init_equilibrium()
        lat.set_parallel(True)
        g = GpuLattice(w, h, True, seed, bands)
        g.set_inlet(8, 0.22, (0.3, 0.0))
        g.load(lat.cells)
        g.set_resident(True)
        runs.append((lat, g))
    for rep in range(3):
        for _, g in runs:
            g.advance_async(37)
        for lat, g in runs:
            lat.advance(37)
            g.wait()
            assert np.array_equal(g.store(), lat.cells), f"{lat.w}x{lat.h} after {lat.steps} steps"
    for _, g in runs:
        g.close()


def test_field_reattached_and_restart_between_resident_calls(oracle):
    """Resident launches that coarse-grain in the launch across a field
    re-attached with other blocks (more block rows: the arrival counts and
    the block sums are reallocated) and across a restart with new cells and
    seed: cells and field bits equal the oracle's throughout."""
    w, h, bands = 512, 300, 6
    say(f"resident reattach/restart {w}x{h}")
    lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, bands, 0xFEED)
    lat.init_equilibrium()
    g = GpuLattice(w, h, True, 0xFEED, bands)
    g.set_inlet(8, 0.22, (0.3, 0.0))
    g.load(lat.cells)
    g.set_resident(True)

    def check(f):
        assert np.array_equal(g.store(), lat.cells), f"cells after {lat.steps} steps"
        for a, b in zip(g.read_field(), f.read()):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"field after {lat.steps}"

    for bx, by in ((16, 16), (8, 4), (40, 30)):
        g.attach_field(bx, by)
        f = oracle.Field(w, h, bx, by)
        for n, every in ((30, 10), (17, 40)):
            replay_sampling(lat, f, n, every, 0.5)
            g.advance_sampling(n, every, 0.5)
            check(f)
    # restart: new cells (the oracle's after more steps) and a new seed
    lat2 = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, bands, 0xBEEF)
    lat2.init_equilibrium()
    lat2.advance(7)
    cells = lat2.cells.copy()
    lat.restart(cells, 0xABC)
    g.restart(cells, 0xABC)
    f = oracle.Field(w, h, 40, 30)
    g.attach_field(40, 30)
    replay_sampling(lat, f, 25, 10, 0.5)
    g.advance_sampling(25, 10, 0.5)
    check(f)
    g.close()
