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


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
    config.addinivalue_line("markers", "slow: longer CPU replays")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O
