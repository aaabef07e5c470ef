"""Writes rng_kat.json: known answers of the CPU random streams, from the
oracle (which reproduces all five GOLDEN rows of tests/golden.rs:141-147
bit for bit, so these are the reference's numbers).

    python3 tests/golden/make_rng_kat.py

Rng::new(seed) states and first next_u64 outputs (src/rng.rs:8-29), and the
seeds of the per-band collision streams and of the inflow stream for a few
steps (src/lattice.rs:233,242,276)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

STEP_MIX = 0x9E3779B97F4A7C15
INLET_MIX = 0xD1B54A32D192ED03
M64 = (1 << 64) - 1


def main():
    O.build()
    seeds = [0, 1, 42, 7, 0x5EED, 0xC0FFEE, M64]
    kat = {"rng": [{"seed": s, "state": O.rng_state(s),
                    "next_u64": [int(v) for v in O.rng_u64(s, 8)]} for s in seeds],
           "streams": []}
    for seed in (42, 0x5EED):
        for t in (0, 1, 99):
            bands = []
            for b in range(3):
                base = (seed ^ ((t * STEP_MIX) & M64)) ^ ((b + 1) << 32)
                bands.append({"band": b, "seed": base, "next_u64": [int(v) for v in O.rng_u64(base, 4)]})
            inlet = seed ^ (((t + 1) * INLET_MIX) & M64)
            kat["streams"].append({"seed": seed, "step": t, "bands": bands,
                                   "inlet_seed": inlet,
                                   "inlet_next_u64": [int(v) for v in O.rng_u64(inlet, 4)]})
    with open(os.path.join(HERE, "rng_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
