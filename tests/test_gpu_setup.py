"""Device-side set-up against the oracle, and the Sim facade, through the C
ABI (gpu): the reference's cylinder obstacle (src/main.rs:589-592) built on
the device by lg_add_solid_disc, and lgca_amd.Sim driven as the reference's
Sim (src/sim.rs:35-174) is, each bit for bit with the CPU path. Each case
prints what it is first."""
import numpy as np
import pytest

from lgca_amd import GpuLattice, Sim

pytestmark = pytest.mark.gpu

SQRT3_2 = np.float32(0.8660254)


def f32(x):
    return np.float32(x)


def cylinder(w, h, size, x):
    """src/main.rs:582-592: r = size / 2, centre (obstacle_x, h sqrt3/2 / 2)."""
    yc = f32(f32(h) * SQRT3_2 / f32(2.0))
    return float(f32(x)), float(yc), float(f32(f32(size) / f32(2.0)))


@pytest.mark.parametrize("w,h,size,x,rest,bands", [
    (512, 256, 64.0, 128.0, True, 4),     # the reference's default shape, scaled
    (200, 90, 37.5, 61.3, True, 3),       # ragged width, off-grid centre
    (130, 64, 9.0, 3.0, False, 1),        # a disc across the periodic edge's start
    (4096, 2304, 288.0 * 0.8660254, 1024.0, True, 16),  # the headline lattice's scale
])
def test_device_disc_equals_the_cpu_cylinder(oracle, w, h, size, x, rest, bands):
    cx, cy, r = cylinder(w, h, size, x)
    print(f"disc {w}x{h} centre ({cx}, {cy}) r {r}", flush=True)
    lat = oracle.Lattice(w, h, 0.25, (0.2, 0.0), 8, rest, bands, 0xD15C + w)
    lat.add_solid_disc(cx, cy, r)
    lat.init_equilibrium()
    g = GpuLattice(w, h, rest, 0xD15C + w, bands)
    g.set_inlet(8, 0.25, (0.2, 0.0))
    g.add_solid_disc(cx, cy, r)
    g.init_equilibrium(0.25, 0.2, 0.0)
    got = g.store()
    assert np.count_nonzero(got & 0x80) == np.count_nonzero(lat.cells & 0x80) > 0
    assert np.array_equal(got, lat.cells)
    lat.advance(30)
    g.advance(30)
    assert np.array_equal(g.store(), lat.cells)
    g.close()


def test_sim_facade_follows_the_reference_sim(oracle):
    """Sim::new from a prepared lattice, a sample after warmup, then
    advance_sampling in uneven calls (samples every k and at each call's
    end), advance, total_particles, cells and restart -- against the same
    calls on the oracle (src/sim.rs:119-174)."""
    w, h, bx, alpha, every = 256, 128, 16, 0.25, 7
    print("sim facade", flush=True)
    lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, 4, 0x5EED)
    cx, cy, r = cylinder(w, h, 32.0, 64.0)
    lat.add_solid_disc(cx, cy, r)
    lat.init_equilibrium()
    start = lat.cells.copy()
    sim = Sim(w, h, True, 0x5EED, 4, 8, 0.22, (0.3, 0.0), start)
    sim.attach_field(bx, bx)
    field = oracle.Field(w, h, bx, bx)
    lat.advance(20)
    sim.advance(20)
    field.sample(lat, 1.0)
    got = sim.sample(1.0)
    for a, b in zip(got, field.read()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for n in (30, 12, 5):
        left = n
        while left > 0:
            take = min(every, left)
            lat.advance(take)
            field.sample(lat, alpha)
            left -= take
        got = sim.advance_sampling(n, every, alpha)
        for a, b in zip(got, field.read()):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(sim.cells(), lat.cells)
    assert sim.total_particles() == lat.total_particles()
    # restart: the initial cells with a new seed, counted from step 0
    lat2 = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, 4, 0x5EED)
    lat2.restart(start, 77)
    sim.restart(start, 77)
    lat2.advance(25)
    sim.advance(25)
    assert np.array_equal(sim.cells(), lat2.cells)


def test_compiled_header_consumer_runs_the_golden_case():
    """tests/harness/sim_harness.cpp: a C++ program that includes
    include/lgca.h and links the library, drives a golden channel/plate case as
    the reference's Sim does (src/sim.rs:35-174) and checks the cells' and the
    averaged field's hashes, the particle count and a short-buffer refusal."""
    import subprocess
    from conftest import build_harness, HARNESS_BIN
    import os
    exe = HARNESS_BIN if os.path.exists(HARNESS_BIN) else build_harness()
    print(f"harness {exe}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "harness: ok" in r.stdout


def test_memory_cache_reuses_and_trims(oracle):
    """Buffers of destroyed lattices go to the library's cache and come back to
    the next lattices (whose runs stay exact); lg_trim_memory gives the cache
    back to the runtime, and lattices made after it still run exactly."""
    import lgca_amd
    lgca_amd.trim_memory()
    for rnd in range(3):
        gs = []
        for k in range(4):
            w, h = 256 + 64 * k, 128
            lat = oracle.Lattice(w, h, 0.22, (0.3, 0.0), 8, True, 2, 90 + k)
            lat.init_equilibrium()
            g = GpuLattice(w, h, True, 90 + k, 2)
            g.set_inlet(8, 0.22, (0.3, 0.0))
            g.load(lat.cells)
            gs.append((lat, g))
        for lat, g in gs:
            lat.advance(23)
            g.advance(23)
            assert np.array_equal(g.store(), lat.cells), f"round {rnd}, {lat.w} x {lat.h}"
        for _, g in gs:
            g.close()
        freed = lgca_amd.trim_memory() if rnd == 1 else None
        if rnd == 1:
            assert freed > 0
    assert lgca_amd.trim_memory() > 0
    assert lgca_amd.trim_memory() == 0
