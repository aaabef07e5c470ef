"""The C-ABI library loads and exports every symbol include/lgca.h declares
(no device calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

import lgca_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(lg_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_gpulattice_methods():
    names = declared_functions()
    # one entry point per GpuLattice method (src/gpu.rs:478-755)
    for n in ("lg_create", "lg_destroy", "lg_device_name", "lg_set_inlet", "lg_attach_field",
              "lg_load", "lg_restart", "lg_store", "lg_advance", "lg_advance_sampling",
              "lg_sample_field", "lg_read_field", "lg_total_particles", "lg_steps"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lgca_amd.LIB_PATH):
        pytest.skip("liblgca_hip.so not built (run __graft_entry__.build())")
    lib = C.CDLL(lgca_amd.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    lgca_amd.load_library()  # declares every signature the Python side uses
