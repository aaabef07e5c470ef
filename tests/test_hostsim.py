"""The device algorithm (lattice-gas_amd/csrc/lgca_device.h) compiled for the
host and compared with the oracle, cell for cell, every step. This checks the
GPU path's logic -- propagation shifts, parity, wrap, ragged tails, jump
offsets into the band streams, the generated circuit, inflow order -- on the
CPU; the gpu tests run the same source on the device."""
import ctypes as C

import numpy as np
import pytest

from conftest import build_hostsim


@pytest.fixture(scope="module")
def hs():
    L = C.CDLL(build_hostsim())
    P = C.POINTER
    L.hs_new.restype = C.c_void_p
    L.hs_new.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, P(C.c_uint32)]
    L.hs_free.argtypes = [C.c_void_p]
    L.hs_set_units.argtypes = [C.c_void_p, C.c_uint32]
    L.hs_set_passes.argtypes = [C.c_void_p, C.c_uint32]
    L.hs_set_draw_units.argtypes = [C.c_void_p, C.c_uint32]
    L.hs_load.argtypes = [C.c_void_p, P(C.c_uint8)]
    L.hs_store.argtypes = [C.c_void_p, P(C.c_uint8)]
    L.hs_advance.argtypes = [C.c_void_p, C.c_uint64]
    L.hs_collide_table.argtypes = [C.c_int, P(C.c_uint8)]
    L.hs_charpoly.restype = C.c_uint64
    L.hs_jump_state.restype = C.c_uint64
    L.hs_jump_state.argtypes = [C.c_uint64, C.c_uint64]
    L.hs_mulz_check.restype = C.c_int
    L.hs_mulz_check.argtypes = [C.c_uint64, C.c_int]
    L.hs_step2_roles.argtypes = [C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_uint32)]
    return L


def test_table_multiplication_by_z_powers(hs):
    """The byte-table c(z) z^f mod P used for the jump tables equals the
    bit-serial product."""
    assert hs.hs_mulz_check(12345, 200000) == 0


@pytest.mark.parametrize("cblocks,dblocks", [(576, 576), (576, 1152), (1152, 576), (1, 1), (0, 5),
                                             (5, 0), (7, 9), (13, 8), (8, 8), (17, 16), (100, 3)])
def test_two_pass_block_roles(hs, cblocks, dblocks):
    """The launch's block -> (role, block of that role) map of the two-pass
    step covers every collide and draw block once, alternating roles while
    both have blocks left."""
    n = cblocks + dblocks
    draw = np.zeros(max(n, 1), dtype=np.uint32)
    rb = np.zeros(max(n, 1), dtype=np.uint32)
    hs.hs_step2_roles(cblocks, dblocks, draw.ctypes.data_as(C.POINTER(C.c_uint32)),
                      rb.ctypes.data_as(C.POINTER(C.c_uint32)))
    draw, rb = draw[:n], rb[:n]
    assert sorted(rb[draw == 0]) == list(range(cblocks))
    assert sorted(rb[draw == 1]) == list(range(dblocks))
    paired = 2 * min(cblocks, dblocks)
    assert (draw[1:paired] != draw[:paired - 1]).all() if paired > 1 else True


def u8p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class HostSim:
    def __init__(self, L, w, h, rest, seed, bands, inlet, thresh):
        self.L, self.w, self.h = L, w, h
        t = np.ascontiguousarray(thresh, dtype=np.uint32)
        self.p = L.hs_new(w, h, int(rest), seed, bands, inlet, t.ctypes.data_as(C.POINTER(C.c_uint32)))

    def set_units(self, units):
        self.L.hs_set_units(self.p, units)

    def set_passes(self, passes, draw_units=2):
        self.L.hs_set_passes(self.p, passes)
        self.L.hs_set_draw_units(self.p, draw_units)

    def load(self, cells):
        a = np.ascontiguousarray(cells, dtype=np.uint8)
        self.L.hs_load(self.p, u8p(a))

    def store(self):
        out = np.zeros(self.w * self.h, dtype=np.uint8)
        self.L.hs_store(self.p, u8p(out))
        return out

    def advance(self, n):
        self.L.hs_advance(self.p, n)

    def __del__(self):
        self.L.hs_free(self.p)


@pytest.mark.parametrize("rest", [True, False])
def test_circuit_matches_the_collision_table(oracle, hs, rest):
    stride, table = oracle.collision_table(rest)
    out = np.zeros(table.size, dtype=np.uint8)
    hs.hs_collide_table(int(rest), u8p(out))
    assert np.array_equal(out, table.ravel())


def test_charpoly_is_primitive_degree_64(hs):
    """z^(2^64-1) == 1 mod P and z^((2^64-1)/q) != 1 for the prime factors q of
    2^64-1 (xorshift64's full period), evaluated through jump_state: T^n s = s
    exactly when n is a multiple of the period."""
    s = 0x0123456789ABCDEF
    period = (1 << 64) - 1
    assert hs.hs_charpoly() != 0
    assert hs.hs_jump_state(s, period) == s
    for q in (3, 5, 17, 257, 641, 65537, 6700417):
        assert hs.hs_jump_state(s, period // q) != s


def test_jump_equals_sequential_steps(oracle, hs):
    M = (1 << 64) - 1
    s0 = oracle.rng_state(0xABC)
    s, ns = s0, 0
    for n in (1, 2, 3, 63, 64, 65, 1000, 4096, 20000):
        while ns < n:
            s ^= (s << 13) & M
            s ^= s >> 7
            s ^= (s << 17) & M
            ns += 1
        assert hs.hs_jump_state(s0, n) == s
    # composition for large offsets: T^(a+b) = T^a T^b
    for a, b in ((2**40 + 5, 2**33 + 7), (123456789123, 987654321)):
        assert hs.hs_jump_state(hs.hs_jump_state(s0, a), b) == hs.hs_jump_state(s0, a + b)


def compare(oracle, hs, w, h, rest, bands, inlet, seed, steps, disc=None, box=None,
            density=0.22, inflow=(0.3, 0.0), units=1, passes=(1, 2)):
    lat = oracle.Lattice(w, h, density, inflow, inlet, rest, bands, seed)
    if disc:
        lat.add_solid_disc(*disc)
    if box:
        lat.add_solid_box(*box)
    lat.init_equilibrium()
    sim = HostSim(hs, w, h, rest, seed, bands, inlet, oracle.thresholds(density, *inflow, rest))
    sim.set_units(units)
    sim.set_passes(*passes)
    sim.load(lat.cells)
    for s in range(steps):
        lat.advance(1)
        sim.advance(1)
        got, ref = sim.store(), lat.cells
        if not np.array_equal(got, ref):
            bad = np.nonzero(got != ref)[0]
            pytest.fail(f"step {s + 1}: {bad.size} cells differ, first at "
                        f"{[(int(b) % w, int(b) // w) for b in bad[:4]]}")


PIPELINES = [(1, 2), (2, 2), (2, 1)]  # (passes, draw units per word)


@pytest.mark.parametrize("passes", PIPELINES)
@pytest.mark.parametrize("units", [1, 2])
@pytest.mark.parametrize("idx", range(5))
def test_golden_cases_every_step(oracle, hs, idx, units, passes):
    name, w, h, rest, inlet, plate, threads, steps = oracle.GOLDEN_CASES[idx]
    box = oracle.golden_plate(w, h) if plate else None
    compare(oracle, hs, w, h, rest, threads, inlet, oracle.GOLDEN_SEED, 60, box=box, units=units,
            passes=passes)


@pytest.mark.parametrize("w,h,rest,bands,inlet", [
    (65, 6, True, 1, 8), (100, 18, True, 3, 8), (129, 64, False, 4, 70), (200, 40, True, 7, 130),
    (64, 32, True, 5, 0), (320, 10, True, 10, 8), (80, 48, True, 2, 80), (127, 8, False, 3, 127),
])
@pytest.mark.parametrize("passes", PIPELINES)
@pytest.mark.parametrize("units", [1, 2])
def test_ragged_shapes_every_step(oracle, hs, w, h, rest, bands, inlet, units, passes):
    yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
    compare(oracle, hs, w, h, rest, bands, inlet, 0xC0FFEE + w, 25, disc=(w / 3.0, yc, 3.5),
            units=units, passes=passes)


@pytest.mark.parametrize("w,h,rest,bands,inlet", [
    (128, 64, True, 4, 8), (65, 6, True, 1, 8), (100, 18, True, 3, 8), (129, 40, False, 4, 70),
    (200, 40, True, 7, 130), (64, 32, True, 5, 0), (127, 8, False, 3, 127),
])
@pytest.mark.parametrize("ks,L", [(2, 6), (2, 14), (4, 18), (8, 32), (3, 7), (8, 17)])
def test_temporal_blocking_every_group(oracle, hs, w, h, rest, bands, inlet, ks, L):
    """The temporally blocked collide (tiles of L rows stepped ks steps with
    2 ks halo rows, only the middle rows kept) gives the oracle's cells after
    every group of ks steps, including tiles taller than the lattice and a
    last tile that wraps past the last row."""
    hs.hs_advance_tiled.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32]
    density, inflow, seed = 0.22, (0.3, 0.0), 0xB10C + w
    lat = oracle.Lattice(w, h, density, inflow, inlet, rest, bands, seed)
    yc = float(np.float32(np.float32(h) * oracle.SQRT3_2 / np.float32(2)))
    lat.add_solid_disc(w / 3.0, yc, 3.5)
    lat.init_equilibrium()
    sim = HostSim(hs, w, h, rest, seed, bands, inlet, oracle.thresholds(density, *inflow, rest))
    sim.load(lat.cells)
    for g in range(3):
        lat.advance(ks)
        hs.hs_advance_tiled(sim.p, ks, ks, L)
        got, ref = sim.store(), lat.cells
        if not np.array_equal(got, ref):
            bad = np.nonzero(got != ref)[0]
            pytest.fail(f"group {g + 1}: {bad.size} cells differ, first at "
                        f"{[(int(b) % w, int(b) // w) for b in bad[:4]]}")
    # a ragged final group (n not a multiple of ks)
    lat.advance(ks + 1)
    hs.hs_advance_tiled(sim.p, ks + 1, ks, L)
    assert np.array_equal(sim.store(), lat.cells)


def test_branch_free_masked_draw(hs):
    """The resident bands' branch-free masked half-word draw equals the
    branching one (digits and stream state) for random and full masks."""
    assert hs.hs_draw_sel_check(0x5E1, 20000) == 0


def test_single_run_draw(hs):
    """The resident bands' plain-chain draw for a half whose non-drawing cells
    form one run (digits moved past the run) equals the branching masked
    draw, for runs anywhere and of any length; lg_one_run says whether a
    mask's zero bits form at most one run."""
    hs.hs_draw_run_check.argtypes = [C.c_uint64, C.c_int]
    assert hs.hs_draw_run_check(0x2A11, 20000) == 0


@pytest.mark.parametrize("w,h", [(4096, 64), (300, 40), (100, 30), (64, 16), (129, 20)])
def test_lead_jumps(hs, w, h):
    """Jumps that start a word's chain as many draws early as its low half
    has cells without a draw (one run: box chords, the row's tail): the
    plain chain's digits moved into place equal the masked draw's, and the
    chain ends where the high half's own jump starts."""
    hs.hs_lead_check.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
    assert hs.hs_lead_check(w, h, 0x1EAD + w) == 0


def test_pow4_windows_reach_group_starts(hs):
    """State 4g of a stream through the T^(4g) windows equals 4g sequential
    xorshift steps, and tables assembled group by group from those states
    equal the tables built from the first 64 states."""
    hs.hs_pow4_check.argtypes = [C.c_uint64, C.c_int]
    assert hs.hs_pow4_check(0x9047, 500) == 0


@pytest.mark.parametrize("w,h,cols,rest", [(4096, 64, 8, 1), (100, 40, 64, 1), (65, 30, 65, 0),
                                            (300, 50, 37, 1), (64, 10, 1, 0)])
def test_per_cell_inflow_jumps(hs, w, h, cols, rest):
    """Per-cell inflow jumps give every fluid inflow cell the same sample as
    the serial per-word inflow chain, with solid cells taking no draw."""
    hs.hs_inflow_cells_check.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint64]
    assert hs.hs_inflow_cells_check(w, h, min(cols, 64), rest, 0xF10 + w) == 0
