"""The oracle's building blocks against what the mathematics fixes.

Each check is chosen so a plausible slip (a dropped term, a wrong sign, a
transposed table) fails it: conservation and closure of the collision table
(src/collision.rs:126-150), the class census recomputed from scratch, the
moment tables against a bit-by-bit count (src/moments.rs:147-211), the
equilibrium sampler's statistics (src/lattice.rs:471-559), straight flight
and wall reversal (src/lattice.rs:408-465), and threshold edge cases.
"""
import itertools
import math

import numpy as np
import pytest

CX2 = [2, 1, -1, -2, -1, 1]
CY2 = [0, 1, 1, 0, -1, -1]
CXF = [1.0, 0.5, -0.5, -1.0, -0.5, 0.5]
S = 0.8660254
CYF = [0.0, S, S, 0.0, -S, -S]


def inv(s):
    m = sum((s >> d) & 1 for d in range(7))
    return (m, sum(CX2[d] for d in range(6) if s >> d & 1), sum(CY2[d] for d in range(6) if s >> d & 1))


@pytest.mark.parametrize("rest", [False, True])
def test_collisions_conserve_mass_and_momentum(oracle, rest):
    stride, t = oracle.collision_table(rest)
    assert stride == (30 if rest else 6)
    for s in range(t.shape[0]):
        for k in range(stride):
            assert inv(s) == inv(int(t[s, k]))


def test_groups_are_closed(oracle):
    stride, t = oracle.collision_table(True)
    for s in range(128):
        for k in range(stride):
            assert s in t[int(t[s, k])]


@pytest.mark.parametrize("rest,n_classes,sizes", [
    (True, 80, {1: 52, 2: 12, 3: 14, 5: 2}),
    (False, 53, {1: 44, 2: 7, 3: 2}),
])
def test_class_census(oracle, rest, n_classes, sizes):
    """Recomputed here by brute force over states, independent of the oracle."""
    groups = {}
    for s in range(128 if rest else 64):
        groups.setdefault(inv(s), []).append(s)
    cl = oracle.classes(rest)
    assert len(cl) == len(groups) == n_classes
    census = {}
    for g in cl:
        census[len(g)] = census.get(len(g), 0) + 1
    assert census == sizes
    # order: by (mass, px2, py2), members ascending (src/collision.rs:62-69)
    assert cl == [groups[k] for k in sorted(groups)]


@pytest.mark.parametrize("rest", [False, True])
def test_table_is_uniform_over_each_class(oracle, rest):
    """table[s][k] = g[k % |g|]: each member appears stride/|g| times per row."""
    stride, t = oracle.collision_table(rest)
    for g in oracle.classes(rest):
        for s in g:
            row = list(t[s])
            for m in g:
                assert row.count(m) == stride // len(g)
            assert row == [g[k % len(g)] for k in range(stride)]


def test_moment_tables_match_bit_by_bit(oracle):
    fluid, walls = oracle.moment_tables()
    for c in range(256):
        solid = bool(c & 0x80)
        for tab, count_walls in ((fluid, False), (walls, True)):
            v = int(tab[c])
            m, px, py = inv(c & 0x7F) if (count_walls or not solid) else (0, 0, 0)
            assert v & 0xFFFF == m
            assert ((v >> 16) & 0xFFFF) - 4 == px
            assert ((v >> 32) & 0xFFFF) - 2 == py
            assert (v >> 48) & 0xFFFF == int(solid)


def test_a_full_block_does_not_overflow(oracle):
    for c in (0, 0x3F, 0x7F, 0x41, 0xFF):
        cells = np.full(8191, c, dtype=np.uint8)
        (m, px, py, s), _ = oracle.moments_sum(cells, True)
        m1, px1, py1 = inv(c & 0x7F)
        assert (m, px, py, s) == (8191 * m1, 8191 * px1, 8191 * py1, 8191 * int(c >= 0x80))


def test_chunking_is_invisible(oracle):
    x, cells = 1, []
    for _ in range(8191 * 3 + 17):
        x = (x * 1664525 + 1013904223) & 0xFFFFFFFF
        cells.append(x >> 24)
    cells = np.array(cells, dtype=np.uint8)
    whole, _ = oracle.moments_sum(cells, False)
    parts = [oracle.moments_sum(cells[i:i + 1000], False)[0] for i in range(0, cells.size, 1000)]
    assert whole == tuple(sum(p[i] for p in parts) for i in range(4))


def test_rng_is_textbook_xorshift64star(oracle):
    """src/rng.rs: splitmix64 finaliser on seed ^ golden ratio, odd; then
    xorshift (13, 7, 17) with the 0x2545F4914F6CDD1D output multiplier."""
    M = (1 << 64) - 1
    for seed in (0, 1, 42, 0x5EED, 2**64 - 1):
        s = seed ^ 0x9E3779B97F4A7C15
        s = ((s ^ (s >> 30)) * 0xBF58476D1CE4E5B9) & M
        s = ((s ^ (s >> 27)) * 0x94D049BB133111EB) & M
        s = (s ^ (s >> 31)) | 1
        assert oracle.rng_state(seed) == s
        outs = oracle.rng_u64(seed, 5)
        for o in outs:
            s ^= (s << 13) & M
            s ^= s >> 7
            s ^= (s << 17) & M
            assert int(o) == (s * 0x2545F4914F6CDD1D) & M


def test_threshold_edges(oracle):
    # p = 0.25 exactly: ceil(2^22) = 2^22; p >= 1 clamps to 2^24; p <= 0 to 0
    t = oracle.thresholds(0.25, 0.0, 0.0, True)
    assert list(t) == [1 << 22] * 7
    assert list(oracle.thresholds(0.0, 0.0, 0.0, True)) == [0] * 7
    t = oracle.thresholds(0.5, 1.0, 0.0, False)  # east p = 0.5 * 3 -> clamp
    assert t[0] == 1 << 24 and t[3] == 0 and t[6] == 0
    t = oracle.thresholds(float("nan"), 0.0, 0.0, True)
    assert list(t) == [0] * 7


@pytest.mark.parametrize("density,ux,uy", [(0.22, 0.3, 0.0), (0.5, 0.0, 0.2)])
@pytest.mark.parametrize("rest", [False, True])
def test_every_direction_hits_its_requested_rate(oracle, density, ux, uy, rest):
    """src/lattice.rs:471-499: within 5 sigma at 1e6 draws."""
    n = 1_000_000
    s = oracle.eq_samples(density, ux, uy, rest, 0xC0FFEE, n)
    bias = 2.0 * 7.0 / 6.0 if rest else 2.0
    for d in range(6):
        want = density * (1 + bias * (CXF[d] * ux + CYF[d] * uy))
        got = np.count_nonzero(s & (1 << d)) / n
        assert abs(got - want) < 5 * math.sqrt(want * (1 - want) / n)
    if rest:
        got = np.count_nonzero(s & 0x40) / n
        assert abs(got - density) < 5 * math.sqrt(density * (1 - density) / n)
    else:
        assert not (s & 0x40).any()


def test_paired_directions_are_uncorrelated(oracle):
    """src/lattice.rs:503-531."""
    n = 1_000_000
    s = oracle.eq_samples(0.5, 0.0, 0.0, False, 0xBEEF, n)
    bits = [(s >> d) & 1 for d in range(6)]
    for a, b in itertools.combinations(range(6), 2):
        pa, pb = bits[a].mean(), bits[b].mean()
        pab = (bits[a] & bits[b]).mean()
        corr = (pab - pa * pb) / math.sqrt(pa * (1 - pa) * pb * (1 - pb))
        assert abs(corr) < 0.01


@pytest.mark.parametrize("rest", [False, True])
def test_equilibrium_hits_the_requested_velocity(oracle, rest):
    """src/lattice.rs:535-559."""
    s = oracle.eq_samples(0.2, 0.3, 0.0, rest, 99, 200_000)
    px = sum(CXF[d] * np.count_nonzero(s & (1 << d)) for d in range(6))
    mass = sum(np.count_nonzero(s & (1 << d)) for d in range(7))
    assert abs(px / mass - 0.3) < 0.01


def test_single_particle_travels_in_a_straight_line(oracle):
    """src/lattice.rs:408-439."""
    for d in range(6):
        for row in (10, 11):
            lat = oracle.Lattice(48, 48, 0.2, (0.0, 0.0), 0, False, 1, 1)
            cells = np.zeros(48 * 48, dtype=np.uint8)
            cells[row * 48 + 24] = 1 << d
            lat.cells = cells
            lat.advance(8)
            out = lat.cells
            occ = np.nonzero(out & 0x7F)[0]
            assert occ.size == 1 and out[occ[0]] == 1 << d
            x, y = int(occ[0]) % 48, int(occ[0]) // 48
            px = x + (0.5 if y & 1 else 0.0)
            sx = 24 + (0.5 if row & 1 else 0.0)
            assert abs(px - (sx + 8 * CXF[d])) < 1e-3
            assert abs(y * S - (row * S + 8 * CYF[d])) < 1e-3


def test_walls_reverse_particles(oracle):
    """src/lattice.rs:442-465."""
    lat = oracle.Lattice(48, 48, 0.2, (0.0, 0.0), 0, False, 1, 1)
    cells = np.zeros(48 * 48, dtype=np.uint8)
    cells[24 * 48 + 23] = 0x80
    cells[24 * 48 + 20] = 1
    lat.cells = cells
    lat.advance(6)
    out = lat.cells
    occ = np.nonzero(out & 0x7F)[0]
    assert occ.size == 1 and occ[0] == 24 * 48 + 20 and out[occ[0]] & 0x7F == 1 << 3


@pytest.mark.parametrize("rest", [False, True])
def test_periodic_box_conserves(oracle, rest):
    """src/lattice.rs:396-404."""
    lat = oracle.Lattice(32, 32, 0.2, (0.0, 0.0), 0, rest, 1, 12345)
    s = oracle.rng_u64(7, 32 * 32)
    lat.cells = ((s >> 32).astype(np.uint32) & (0x7F if rest else 0x3F)).astype(np.uint8)
    before = oracle.exact_invariants(lat.cells)
    lat.advance(200)
    assert oracle.exact_invariants(lat.cells) == before
