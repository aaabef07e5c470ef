"""The batched-ensemble workload (256 lattices of 256..2048 a side, plates
and inflow) in a batch, through the C ABI (gpu): 32 of its members -- the
two largest, the four smallest and 26 more -- in one batch with its
XCD-segmented tile table, autotuned, then advanced through graph replays
and a remainder, each equal to its own solo run, and the four smallest to
the oracle's replay."""
import numpy as np
import pytest

from lgca_amd import Batch
import workloads as W
from oracle.workloads_cpu import config_oracle

pytestmark = pytest.mark.gpu


def pick(cfgs):
    cells = [c["w"] * c["h"] for c in cfgs]
    order = sorted(range(len(cfgs)), key=lambda i: cells[i])
    chosen = order[:4] + order[-2:]
    chosen += [i for i in range(len(cfgs)) if i not in chosen][:32 - len(chosen)]
    return [cfgs[i] for i in chosen], [cfgs[i] for i in order[:4]]


def test_ensemble_members_equal_their_solo_runs(oracle):
    cfgs, smallest = pick(W.config4(256))
    assert max(max(c["w"], c["h"]) for c in cfgs) == 2048
    print(f"ensemble: {len(cfgs)} members, {sum(c['w'] * c['h'] for c in cfgs)} cells", flush=True)
    members = [W.config_gpu(c) for c in cfgs]
    b = Batch(members)
    used = b.autotune(budget_steps=256)  # two candidates: 32 warm + 3 x 32 timed steps each
    assert used > 0
    tile, grid, chunk = b.shape()
    n = 2 * chunk + 7          # two graph replays and a remainder
    b.advance(n)
    total = used + n
    print(f"ensemble: autotune used {used} steps, shape {b.shape()} pipeline {b.pipeline()}, "
          f"then {n} more", flush=True)
    for c, g in zip(cfgs, members):
        assert g.steps == total
        solo = W.config_gpu(c)
        solo.advance(total)
        assert np.array_equal(g.store(), solo.store()), c["name"]
        solo.close()
    for c in smallest:
        lat = config_oracle(oracle, c)
        lat.advance(total)
        g = members[cfgs.index(c)]
        assert np.array_equal(g.store(), lat.cells), c["name"]
    b.close()
    for g in members:
        g.close()
