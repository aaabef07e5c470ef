"""Known answers of the CPU random streams (tests/golden/rng_kat.json, written
by tests/golden/make_rng_kat.py from the oracle once it reproduced GOLDEN):
the oracle, and the device source's generator and stream seeds compiled for
the host, both give them."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from conftest import build_hostsim

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "rng_kat.json")))


@pytest.fixture(scope="module")
def hs():
    L = C.CDLL(build_hostsim())
    L.hs_rng_state.restype = C.c_uint64
    L.hs_rng_state.argtypes = [C.c_uint64]
    L.hs_rng_u64.argtypes = [C.c_uint64, C.c_size_t, C.POINTER(C.c_uint64)]
    L.hs_band_seed.restype = C.c_uint64
    L.hs_band_seed.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
    L.hs_inlet_seed.restype = C.c_uint64
    L.hs_inlet_seed.argtypes = [C.c_uint64, C.c_uint64]
    return L


def device_u64(hs, seed, n):
    out = np.zeros(n, dtype=np.uint64)
    hs.hs_rng_u64(seed, n, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return [int(v) for v in out]


def test_oracle_generator(oracle):
    for e in KAT["rng"]:
        assert oracle.rng_state(e["seed"]) == e["state"]
        assert [int(v) for v in oracle.rng_u64(e["seed"], len(e["next_u64"]))] == e["next_u64"]


def test_device_generator_and_stream_seeds(hs):
    for e in KAT["rng"]:
        assert hs.hs_rng_state(e["seed"]) == e["state"]
        assert device_u64(hs, e["seed"], len(e["next_u64"])) == e["next_u64"]
    for st in KAT["streams"]:
        for b in st["bands"]:
            assert hs.hs_band_seed(st["seed"], st["step"], b["band"]) == b["seed"]
            assert device_u64(hs, b["seed"], len(b["next_u64"])) == b["next_u64"]
        assert hs.hs_inlet_seed(st["seed"], st["step"]) == st["inlet_seed"]
        assert device_u64(hs, st["inlet_seed"], 4) == st["inlet_next_u64"]
