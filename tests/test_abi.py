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


def test_compiled_consumer_of_the_header_builds():
    """tests/harness/sim_harness.cpp, which drives the engine as the
    reference's Sim does, compiles against include/lgca.h and links against
    the library (it runs in a gpu test)."""
    from conftest import build_harness
    if not os.path.exists(lgca_amd.LIB_PATH):
        pytest.skip("liblgca_hip.so not built (run __graft_entry__.build())")
    assert os.path.exists(build_harness())


_CTYPES = {
    "int": C.c_int, "size_t": C.c_size_t, "uint64_t": C.c_uint64, "uint32_t": C.c_uint32,
    "float": C.c_float, "double": C.c_double, "void": None,
}
_OPAQUE = ("lg_t", "lg_batch_t", "lg_frames_t")


def _ctype(decl):
    """ctypes type of one C declarator's type ("const float* ux" -> POINTER(c_float))."""
    decl = re.sub(r"\bconst\b", "", decl)
    decl = re.sub(r"\[\s*\d*\s*\]", "*", decl)  # an array parameter is a pointer
    base, stars = re.match(r"\s*([A-Za-z_][A-Za-z0-9_]*)", decl).group(1), decl.count("*")
    if stars == 0:
        return _CTYPES[base]
    if base in _OPAQUE:
        return C.c_void_p if stars == 1 else C.POINTER(C.c_void_p)
    if base == "char" and stars == 1:
        return C.c_char_p
    pt = {"uint8_t": C.c_uint8, "int64_t": C.c_int64, "uint64_t": C.c_uint64,
          "uint32_t": C.c_uint32, "float": C.c_float, "double": C.c_double,
          "size_t": C.c_size_t}[base]
    return C.POINTER(pt) if stars == 1 else C.POINTER(C.POINTER(pt))


def declarations():
    text = open(os.path.join(ROOT, "include", "lgca.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#[^\n]*", "", text)
    out = {}
    for m in re.finditer(r"([A-Za-z_][A-Za-z0-9_ \*]*?)\b(lg_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        params = [] if args in ("", "void") else [a for a in args.split(",")]
        out[name] = (_ctype(ret), [_ctype(a) for a in params])
    return out


def test_python_signatures_match_the_header():
    """The ctypes argument and result types lgca_amd declares are the ones
    include/lgca.h declares, function by function (the signature table is
    written by hand beside the header)."""
    if not os.path.exists(lgca_amd.LIB_PATH):
        pytest.skip("liblgca_hip.so not built (run __graft_entry__.build())")
    L = lgca_amd.load_library()
    decl = declarations()
    assert len(decl) > 60
    checked = 0
    for name, (res, args) in decl.items():
        fn = getattr(L, name)
        if fn.argtypes is None:  # not used from Python
            continue
        def same(a, b):
            if a is None or b is None:
                return a is b
            return a == b or (getattr(a, "_type_", 0) == getattr(b, "_type_", 1))
        assert same(fn.restype, res), f"{name}: result {fn.restype} vs {res}"
        assert len(fn.argtypes) == len(args), f"{name}: {len(fn.argtypes)} args vs {len(args)}"
        for i, (a, b) in enumerate(zip(fn.argtypes, args)):
            assert same(a, b), f"{name} argument {i}: {a} vs {b}"
        checked += 1
    assert checked > 60
