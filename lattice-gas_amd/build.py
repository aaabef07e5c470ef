#!/usr/bin/env python3
"""Build the HIP engine for gfx950: generate the collision circuit, then
compile lattice-gas_amd/csrc/*.hip into liblgca_hip.so (in-tree)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblgca_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no fused multiply-add anywhere: the f32 field must round like the CPU's
    "-ffp-contract=off",
    # packed fp32 ops are kept out of every HIP build on these boxes
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    # resolve every symbol at load time, not at a call's first use
    "-Wl,-z,now",
]


HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
              "-Wall", "-Wextra"]


def build(verbose: bool = False) -> str:
    subprocess.run([sys.executable, os.path.join(HERE, "gen_collide.py"),
                    "--out", os.path.join(CSRC, "lgca_collide.gen.h")], check=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    srcs = [os.path.join(CSRC, f) for f in ("lgca_engine.hip", "lgca_batch.hip", "lgca_band.hip", "lgca_probe.hip")]
    # The frame writers are host-only C++: built by the host compiler into an
    # object that hipcc links into the same library.
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    obj = os.path.join(HERE, "_build", "lgca_frames.o")
    r = subprocess.run(["g++"] + HOST_FLAGS + ["-c", "-o", obj, os.path.join(CSRC, "lgca_frames.cpp")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed for lgca_frames.cpp")
    cmd = [hipcc] + HIPCC_FLAGS + ["-o", tmp] + srcs + ["-x", "none", obj, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for lgca_engine.hip")
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
