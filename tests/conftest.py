"""Shared test setup: the `gpu` marker, import paths, and the CPU-side test
libraries (oracle, host build of the device step), compiled with gcc when
missing. The HIP library itself is built by __graft_entry__.build() and is
never compiled from a test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "lattice-gas_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HOSTSIM_SRC = os.path.join(ROOT, "tests", "hostsim", "lgca_hostsim.cpp")
HOSTSIM_LIB = os.path.join(ROOT, "tests", "hostsim", "build", "liblgca_hostsim.so")


def _hostsim_stale():
    if not os.path.exists(HOSTSIM_LIB):
        return True
    built = os.path.getmtime(HOSTSIM_LIB)
    csrc = os.path.join(ROOT, "lattice-gas_amd", "csrc")
    deps = [HOSTSIM_SRC] + [os.path.join(csrc, f) for f in os.listdir(csrc)
                            if f.endswith((".h", ".hpp"))]
    return any(os.path.getmtime(d) > built for d in deps)


def build_hostsim(force=False):
    if force or _hostsim_stale():
        os.makedirs(os.path.dirname(HOSTSIM_LIB), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        "-o", HOSTSIM_LIB, HOSTSIM_SRC], check=True)
    return HOSTSIM_LIB


HARNESS_SRC = os.path.join(ROOT, "tests", "harness", "sim_harness.cpp")
HARNESS_BIN = os.path.join(ROOT, "tests", "harness", "build", "sim_harness")


def build_harness():
    """The C++ consumer of include/lgca.h (host compiler, linked against the
    built liblgca_hip.so); rebuilt when older than its source, the header or
    the library."""
    lib = os.path.join(ROOT, "lattice-gas_amd", "liblgca_hip.so")
    hdr = os.path.join(ROOT, "include", "lgca.h")
    if not os.path.exists(lib):
        return None
    if not os.path.exists(HARNESS_BIN) or any(
            os.path.getmtime(d) > os.path.getmtime(HARNESS_BIN) for d in (HARNESS_SRC, hdr, lib)):
        os.makedirs(os.path.dirname(HARNESS_BIN), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-o", HARNESS_BIN,
                        HARNESS_SRC, "-L", os.path.dirname(lib), "-llgca_hip",
                        "-Wl,-rpath,$ORIGIN/../../../lattice-gas_amd"], check=True)
    return HARNESS_BIN


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: longer CPU replays")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O
