"""A seeded sweep of random configurations against the oracle (gpu): sizes,
band counts, inflow widths, obstacles, densities, pipelines and work units,
blocking and queued advances. Bit-exact cells, block sums and f32 field."""
import sys

import numpy as np
import pytest

from lgca_amd import GpuLattice

pytestmark = pytest.mark.gpu

# (kernels per step, draw cells per lane, draw-ahead steps, collide cells per lane);
# None: the size rule's defaults
PIPELINES = [(2, 64, 0, 64), (2, 32, 0, 64), (2, 64, 0, 32), (2, 32, 0, 32), (1, 64, 0, 64),
             (1, 64, 0, 32), (2, 64, 2, 64), (2, 64, 8, 32), None]


def case(i):
    """Configuration i of the sweep (seeded; the same on every machine)."""
    r = np.random.default_rng(0x5EED0000 + i)
    w = int(r.integers(5, 401))
    h = 2 * int(r.integers(3, 101))
    rest = bool(r.integers(0, 2))
    bands = int(r.integers(1, min(16, h // 2) + 1))
    inlet = min(w, int(r.choice([0, 1, 8, 63, 64, 65, 130, w])))
    density = float(np.float32(r.uniform(0.05, 0.6)))
    drift = float(np.float32(r.uniform(-0.3, 0.3)))
    kind = int(r.integers(0, 3))
    if kind == 1:
        x0 = float(np.float32(r.uniform(0, w * 0.6)))
        y0 = float(np.float32(r.uniform(0, h * 0.6)))
        obstacle = ("box", (x0, x0 + float(np.float32(r.uniform(1, w * 0.3))),
                            y0, y0 + float(np.float32(r.uniform(1, h * 0.3)))))
    elif kind == 2:
        obstacle = ("disc", (float(np.float32(r.uniform(0, w))), float(np.float32(r.uniform(0, h))),
                             float(np.float32(r.uniform(1.0, min(w, h) / 3.0)))))
    else:
        obstacle = None
    pipeline = PIPELINES[int(r.integers(0, len(PIPELINES)))]
    steps = int(r.integers(10, 61))
    bx, by = int(r.integers(1, min(w, 40) + 1)), int(r.integers(1, min(h, 40) + 1))
    queued = bool(r.integers(0, 2))
    seed = int(r.integers(1, 2**63))
    return dict(w=w, h=h, rest=rest, bands=bands, inlet=inlet, density=density, drift=drift,
                obstacle=obstacle, pipeline=pipeline, steps=steps, block=(bx, by), queued=queued,
                seed=seed)


@pytest.mark.parametrize("i", range(48))
def test_random_configuration(oracle, i):
    c = case(i)
    print(f"sweep {i}: {c}", flush=True)
    sys.stdout.flush()
    lat = oracle.Lattice(c["w"], c["h"], c["density"], (c["drift"], 0.0), c["inlet"], c["rest"],
                         c["bands"], c["seed"])
    if c["obstacle"]:
        kind, args = c["obstacle"]
        (lat.add_solid_box if kind == "box" else lat.add_solid_disc)(*args)
    lat.init_equilibrium()
    g = GpuLattice(c["w"], c["h"], c["rest"], c["seed"], c["bands"])
    if c["pipeline"]:
        passes, draw, ahead, units = c["pipeline"]
        g.set_passes(passes)
        g.set_draw_unit(draw)
        g.set_draw_ahead(ahead)
        g.set_work_unit(units)
    g.set_inlet(c["inlet"], c["density"], (c["drift"], 0.0))
    g.load(lat.cells)
    bx, by = c["block"]
    g.attach_field(bx, by)
    f = oracle.Field(c["w"], c["h"], bx, by)
    half = c["steps"] // 2
    for k in (half, c["steps"] - half):      # two calls, sampled after each
        lat.advance(k)
        f.sample(lat, 0.5)
        if c["queued"]:
            g.advance_async(k, k, 0.5)
        else:
            g.advance_sampling(k, k, 0.5)
    got, ref = g.store(), lat.cells
    if not np.array_equal(got, ref):
        bad = np.nonzero(got != ref)[0]
        pytest.fail(f"{bad.size} cells differ, first at {[(int(b) % c['w'], int(b) // c['w']) for b in bad[:4]]}")
    assert g.steps == lat.steps == c["steps"]
    for a, b in zip(g.read_block_sums(), f.block_sums(lat)):
        assert np.array_equal(a, b)
    for a, b in zip(g.read_field(), f.read()):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    g.close()


@pytest.mark.parametrize("j", range(8))
def test_random_batch(oracle, j):
    """Batches of members drawn from the sweep's configurations (their own
    sizes, bands, inflow and obstacles), a random pipeline and tile shape, and
    two advances with random per-member limits: every member against its own
    oracle replay."""
    from lgca_amd import Batch
    r = np.random.default_rng(0xBA7C0000 + j)
    idx = [int(x) for x in r.choice(48, size=int(r.integers(2, 7)), replace=False)]
    passes, units, draw = int(r.integers(1, 3)), int(r.choice([64, 32])), int(r.choice([64, 32]))
    tile, chunk = int(r.choice([128, 256])), int(r.choice([2, 4, 8]))
    print(f"batch {j}: members {idx} passes={passes} units={units} draw={draw} tile={tile} "
          f"chunk={chunk}", flush=True)
    sys.stdout.flush()
    lats, gs = [], []
    for i in idx:
        c = case(i)
        lat = oracle.Lattice(c["w"], c["h"], c["density"], (c["drift"], 0.0), c["inlet"],
                             c["rest"], c["bands"], c["seed"])
        if c["obstacle"]:
            kind, args = c["obstacle"]
            (lat.add_solid_box if kind == "box" else lat.add_solid_disc)(*args)
        lat.init_equilibrium()
        g = GpuLattice(c["w"], c["h"], c["rest"], c["seed"], c["bands"])
        g.set_inlet(c["inlet"], c["density"], (c["drift"], 0.0))
        g.load(lat.cells)
        lats.append(lat)
        gs.append(g)
    b = Batch(gs)
    b.set_passes(passes)
    b.set_units(units)
    b.set_draw_unit(draw)
    b.set_shape(tile, 0, chunk)
    for _ in range(2):
        n = int(r.integers(5, 40))
        limits = [int(x) for x in r.integers(0, n + 5, size=len(gs))]
        for lat, lim in zip(lats, limits):
            if min(n, lim):
                lat.advance(min(n, lim))
        b.advance_limited(n, limits)
    for i, lat, g in zip(idx, lats, gs):
        assert g.steps == lat.steps, i
        assert np.array_equal(g.store(), lat.cells), f"member {i}"
    b.close()
    for g in gs:
        g.close()
