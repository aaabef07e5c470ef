"""Parity of the HIP engine with the oracle, through the C ABI (gpu).

Bit-exact is the bar everywhere: cells, the f32 coarse field and the integer
block sums. Each case prints what it is before touching the GPU.
"""
import sys

import numpy as np
import pytest

from lgca_amd import GpuLattice, collide_table_on_device
import workloads as W
from oracle.workloads_cpu import config_oracle

pytestmark = pytest.mark.gpu

# tests/golden.rs:141-147
GOLDEN = {
    "periodic/rest/t1": (0x6836a7941f22413e, 0x2ef0cd774a7e93be, 25380, 14999, 177),
    "periodic/no-rest/t1": (0x60bca9c9366b2da7, 0x9ebeb21fffe77bc4, 21733, 12964, -34),
    "periodic/rest/t3": (0x6c568bfe63b36041, 0xc0e475819e2481d6, 25380, 14999, 177),
    "channel/plate/inlet/t1": (0x435cc70a5ddcaef6, 0xe8270bd1fb49fb49, 11701, 6836, 52),
    "channel/plate/inlet/t4": (0x151d9c62d8f873cf, 0x4f0354593ef1135d, 11685, 6777, 127),
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("rest", [True, False])
def test_device_circuit_matches_the_table(oracle, rest):
    """The emitted circuit on the GPU, for every state and every k
    (src/gpu.rs:775-820, against CollisionTable rather than classes)."""
    say(f"circuit rest={rest}")
    stride, table = oracle.collision_table(rest)
    got = collide_table_on_device(rest)
    assert got.shape == table.shape
    assert not (got == 0xFF).any(), "lanes of a word disagreed"
    assert np.array_equal(got, table)


@pytest.mark.parametrize("case", [c[0] for c in W.GOLDEN_CASES])
def test_golden_table_on_the_gpu(oracle, case):
    """tests/golden.rs:156-197 run on the engine: the reference's own
    checksums of cells and field, and exact mass and momentum."""
    c = next(c for c in W.GOLDEN_CASES if c[0] == case)
    name, w, h, rest, inlet, plate, threads, steps = c
    say(f"golden {name} {w}x{h} bands={threads}")
    g = W.golden_gpu(c)
    g.attach_field(16, 16)
    g.advance(steps)
    g.sample_field(1.0)
    cells = g.store()
    ux, uy, rho, sol = g.read_field()
    want = GOLDEN[name]
    assert oracle.fnv1a(cells) == want[0]
    fb = np.concatenate([ux, uy, rho, sol]).astype("<f4").tobytes()
    assert oracle.fnv1a(fb) == want[1]
    assert oracle.exact_invariants(cells) == want[2:5]
    assert g.total_particles() == want[2]


@pytest.mark.parametrize("case", [c[0] for c in W.GOLDEN_CASES])
def test_golden_setup_on_the_device_matches_the_cpu(oracle, case):
    """add_solid + init_equilibrium run on the device give the CPU's cells."""
    c = next(c for c in W.GOLDEN_CASES if c[0] == case)
    say(f"setup {c[0]}")
    lat = oracle.golden_build(c)
    g = W.golden_gpu(c)
    assert np.array_equal(g.store(), lat.cells)


def _compare_run(oracle, w, h, rest, bands, inlet, seed, steps, density=0.22, inflow=(0.3, 0.0),
                 disc=None, box=None, check_every=1, cells_per_lane=64, passes=(2, 32)):
    say(f"run {w}x{h} rest={rest} bands={bands} inlet={inlet} steps={steps} disc={disc} box={box} "
        f"cells/lane={cells_per_lane} passes={passes}")
    lat = oracle.Lattice(w, h, density, inflow, inlet, rest, bands, seed)
    if disc:
        lat.add_solid_disc(*disc)
    if box:
        lat.add_solid_box(*box)
    lat.init_equilibrium()
    g = GpuLattice(w, h, rest, seed, bands)
    g.set_work_unit(cells_per_lane)
    if passes is not None:  # None: the default pipeline, by size
        g.set_passes(passes[0])
        g.set_draw_unit(passes[1])
        g.set_draw_ahead(passes[2] if len(passes) > 2 else 0)
    g.set_inlet(inlet, density, inflow)
    g.load(lat.cells)
    done = 0
    while done < steps:
        k = min(check_every, steps - done)
        lat.advance(k)
        g.advance(k)
        done += k
        got, ref = g.store(), lat.cells
        if not np.array_equal(got, ref):
            bad = np.nonzero(got != ref)[0]
            pytest.fail(f"step {done}: {bad.size} cells differ, first at "
                        f"{[(int(b) % w, int(b) // w) for b in bad[:4]]}")
    assert g.steps == steps == lat.steps
    return lat, g


# (kernels per step, draw cells per lane[, draw-ahead steps])
PIPELINES = [(1, 32), (2, 32), (2, 64), (2, 64, 4), (2, 32, 3)]


@pytest.mark.parametrize("w,h,rest,bands,inlet", [
    (65, 6, True, 1, 8),        # one cell past a word, tiny height
    (100, 18, True, 3, 8),      # ragged tail word, bands of 6
    (129, 64, False, 4, 70),    # inflow wider than a word, no rest
    (200, 40, True, 7, 130),    # bands that do not divide h, 3 inflow words
    (64, 32, True, 5, 0),       # exactly one word per row
    (320, 10, True, 10, 8),     # one row per band
    (80, 48, True, 2, 80),      # inflow over the whole width
])
@pytest.mark.parametrize("passes", PIPELINES)
@pytest.mark.parametrize("cells_per_lane", [64, 32])
def test_ragged_shapes_every_step(oracle, w, h, rest, bands, inlet, cells_per_lane, passes):
    yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
    disc = (w / 3.0, yc, 3.5)
    _compare_run(oracle, w, h, rest, bands, inlet, 0xC0FFEE + w, 40, disc=disc,
                 cells_per_lane=cells_per_lane, passes=passes)


@pytest.mark.parametrize("bands", [1, 2, 3, 5, 8, 16])
def test_band_counts(oracle, bands):
    _compare_run(oracle, 192, 96, True, bands, 8, 0x5EED, 60, check_every=20)


def test_config1_cpu_path_workload(oracle):
    """512x256, rest, density 0.25, drift 0.2 inflow over 8 columns, plate 64
    rows tall at x=128, seed 42, 2000 steps, 16x16 block sums."""
    cfg = W.config1()
    say(f"config1 {cfg}")
    lat = config_oracle(oracle, cfg)
    g = W.config_gpu(cfg)
    assert np.array_equal(g.store(), lat.cells)
    g.attach_field(16, 16)
    lat.advance(cfg["steps"])
    g.advance(cfg["steps"])
    assert np.array_equal(g.store(), lat.cells)
    f = oracle.Field(cfg["w"], cfg["h"], 16, 16)
    f.sample(lat, 1.0)
    g.sample_field(1.0)
    for a, b in zip(g.read_block_sums(), f.block_sums(lat)):
        assert np.array_equal(a, b)
    for a, b in zip(g.read_field(), f.read()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_advance_sampling_cadence_and_running_average(oracle):
    """Samples every k steps and after a ragged last chunk, alpha-blended
    (src/sim.rs:126-136, src/gpu.rs:688-699, src/render.rs:107-115)."""
    w, h, every, alpha = 160, 64, 7, 0.3
    say("advance_sampling")
    lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, 2, 99)
    lat.add_solid_disc(50.0, 27.0, 6.0)
    lat.init_equilibrium()
    g = GpuLattice(w, h, True, 99, 2)
    g.set_inlet(8, 0.22, (0.3, 0.0))
    g.load(lat.cells)
    g.attach_field(20, 16)
    f = oracle.Field(w, h, 20, 16)
    for n in (30, 12, 5):
        left = n
        while left > 0:
            take = min(every, left)
            lat.advance(take)
            f.sample(lat, alpha)
            left -= take
        g.advance_sampling(n, every, alpha)
        for a, b in zip(g.read_field(), f.read()):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_restart_and_resume(oracle):
    """restart resets steps (src/gpu.rs:616-620); set_steps resumes exactly."""
    w, h = 128, 32
    say("restart/resume")
    lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, 3, 5)
    lat.init_equilibrium()
    start = lat.cells
    lat.advance(25)
    mid = lat.cells
    lat.advance(25)
    g = GpuLattice(w, h, True, 123, 3)
    g.set_inlet(8, 0.22, (0.3, 0.0))
    g.restart(start, 5)
    g.advance(50)
    assert np.array_equal(g.store(), lat.cells)
    g2 = GpuLattice(w, h, True, 5, 3)
    g2.set_inlet(8, 0.22, (0.3, 0.0))
    g2.load(mid)
    g2.steps = 25
    g2.advance(25)
    assert np.array_equal(g2.store(), lat.cells)


def test_load_store_round_trip():
    """src/gpu.rs:928-948 at w in {64, 100}, any byte pattern."""
    rng = np.random.default_rng(3)
    for w in (64, 100):
        say(f"roundtrip {w}")
        cells = rng.integers(0, 256, size=w * 20, dtype=np.uint8)
        g = GpuLattice(w, 20, True, 1, 1)
        g.load(cells)
        assert np.array_equal(g.store(), cells)


@pytest.mark.parametrize("w", [48, 80, 130])
def test_single_particle_travels_in_a_straight_line(w):
    """src/lattice.rs:408-439 / src/gpu.rs:825-863 on the engine."""
    cxf = [1.0, 0.5, -0.5, -1.0, -0.5, 0.5]
    cyf = [0.0, 0.8660254, 0.8660254, 0.0, -0.8660254, -0.8660254]
    for d in range(6):
        for row in (10, 11):
            say(f"flight w={w} d={d} row={row}")
            h = 48
            cells = np.zeros(w * h, dtype=np.uint8)
            x0 = w // 2
            cells[row * w + x0] = 1 << d
            g = GpuLattice(w, h, False, 1, 1)
            g.load(cells)
            g.advance(8)
            out = g.store()
            occ = np.nonzero(out & 0x7F)[0]
            assert occ.size == 1 and out[occ[0]] == 1 << d
            x, y = int(occ[0]) % w, int(occ[0]) // w
            px = x + (0.5 if y & 1 else 0.0)
            sx = x0 + (0.5 if row & 1 else 0.0)
            assert abs(px - (sx + 8 * cxf[d])) < 1e-3
            assert abs(y * 0.8660254 - (row * 0.8660254 + 8 * cyf[d])) < 1e-3


def test_walls_reverse_particles():
    """src/lattice.rs:442-465 on the engine."""
    w = h = 48
    say("walls")
    cells = np.zeros(w * h, dtype=np.uint8)
    cells[24 * w + 23] = 0x80
    cells[24 * w + 20] = 1
    g = GpuLattice(w, h, False, 1, 1)
    g.load(cells)
    g.advance(6)
    out = g.store()
    occ = np.nonzero(out & 0x7F)[0]
    assert occ.size == 1 and occ[0] == 24 * w + 20 and out[occ[0]] & 0x7F == 1 << 3


def test_closed_box_conserves_at_full_config2_size():
    """9.4M cells, periodic, no inflow: mass and momentum exact over steps
    (tests/golden.rs:262-297 at the engine's full size)."""
    say("conservation 4096x2304")
    w, h = 4096, 2304
    g = GpuLattice(w, h, True, 42, 16)
    g.init_equilibrium(0.25, 0.2, 0.0)
    c0 = g.store()
    inv0 = W.exact_invariants(c0)
    n0 = g.total_particles()
    g.advance(200)
    c1 = g.store()
    assert W.exact_invariants(c1) == inv0
    assert g.total_particles() == n0 == inv0[0]


@pytest.mark.parametrize("passes", PIPELINES)
@pytest.mark.parametrize("cells_per_lane", [64, 32])
def test_config2_size_matches_the_cpu_replay(oracle, cells_per_lane, passes):
    """Config 2 geometry (4096x2304, plate H/8 at W/4, 8 inflow columns,
    bands 16) for 60 steps against the oracle, every 20 steps."""
    cfg = W.config2(bands=16)
    say(f"config2 replay {cfg} cells/lane={cells_per_lane} passes={passes}")
    lat = config_oracle(oracle, cfg)
    lat.set_parallel(True)
    g = W.config_gpu(cfg)
    g.set_work_unit(cells_per_lane)
    g.set_passes(passes[0])
    g.set_draw_unit(passes[1])
    g.set_draw_ahead(passes[2] if len(passes) > 2 else 0)
    for _ in range(3):
        lat.advance(20)
        g.advance(20)
        assert np.array_equal(g.store(), lat.cells)


@pytest.mark.parametrize("ahead", [1, 2, 8, 64])
def test_draw_ahead_sampling_and_mode_changes(oracle, ahead):
    """Draw-ahead across chunk and sampling boundaries (every 7 steps), with
    the pipeline switched between calls, against the oracle's cells and
    exact block sums."""
    say(f"draw-ahead {ahead}")
    w, h = 320, 96
    lat = oracle.Lattice(w, h, 0.25, (0.2, 0.0), 8, True, 5, 0xABCD)
    lat.add_solid_box(100.0, 104.0, 30.0, 60.0)
    lat.init_equilibrium()
    g = GpuLattice(w, h, True, 0xABCD, 5)
    g.set_inlet(8, 0.25, (0.2, 0.0))
    g.load(lat.cells)
    g.attach_field(16, 16)
    field = oracle.Field(w, h, 16, 16)
    for rep, (mode, n) in enumerate(((ahead, 45), (0, 13), (ahead, 30), (ahead, 1))):
        g.set_draw_ahead(mode)
        g.advance_sampling(n, 7, 1.0)
        done = 0
        while done < n:
            k = min(7, n - done)
            lat.advance(k)
            done += k
        field.sample(lat, 1.0)
        assert np.array_equal(g.store(), lat.cells), f"call {rep}"
        assert [list(x) for x in g.read_block_sums()] == [list(x) for x in field.block_sums(lat)]


def test_default_pipeline_follows_the_lattice_size():
    """passes = 0 (the default) is the two-pass step for the 4096 x 2304 plate
    lattice and the fused step for the 32768 x 16384 shear lattice; an explicit
    choice overrides it."""
    say("default pipeline by size")
    small = GpuLattice(4096, 2304, True, 1, 16)
    assert small.pipeline()[0] == 2
    small.set_passes(1)
    assert small.pipeline()[0] == 1
    small.close()
    big = GpuLattice(32768, 16384, True, 1, 16)
    assert big.pipeline()[0] == 1
    big.set_passes(2)
    assert big.pipeline()[:3] == (2, 64, 64)
    big.close()


def test_default_draw_work_of_small_lattices(oracle):
    """Below 2^16 words: 32-cell draw lanes without inflow, draws 8 steps
    ahead with inflow; explicit settings override, set_passes(0) restores the
    size rule; the default runs equal the CPU replay."""
    say("small-lattice draw defaults")
    g = GpuLattice(2048, 1280, True, 3, 4)
    assert g.pipeline() == (2, 64, 32, 0)
    g.set_inlet(8, 0.22, (0.3, 0.0))
    assert g.pipeline() == (2, 64, 64, 8)
    g.set_draw_unit(32)
    g.set_draw_ahead(0)
    assert g.pipeline() == (2, 64, 32, 0)
    g.set_passes(0)
    assert g.pipeline() == (2, 64, 64, 8)
    g.set_draw_unit(0)
    assert g.pipeline() == (2, 64, 64, 8)
    g.close()
    big = GpuLattice(4096, 2304, True, 3, 4)
    big.set_inlet(8, 0.22, (0.3, 0.0))
    assert big.pipeline() == (2, 64, 64, 0)
    big.close()
    for inlet in (0, 8):
        _compare_run(oracle, 320, 96, True, 3, inlet, 17, 40, disc=(100.0, 40.0, 9.0),
                     passes=None)


@pytest.mark.parametrize("w,h", [(256, 64), (100, 32), (129, 18)])
def test_particle_count_at_ragged_sizes(oracle, w, h):
    """GpuLattice::total_particles at the reference's own test sizes
    (src/gpu.rs:1035-1053), with an obstacle and inflow, after 37 steps:
    equal to the CPU count (src/lattice.rs:292-300)."""
    say(f"particle count {w}x{h}")
    lat, g = _compare_run(oracle, w, h, True, 2, 8, 11, 37, disc=(w / 3.0, h / 3.0, 4.0),
                          passes=None, check_every=37)
    assert g.total_particles() == lat.total_particles()
