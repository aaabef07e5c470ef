"""Continuous batching keeps every served run bit-exact: each result is the
same alone and in the mix, snapshot by snapshot, and equals its CPU replay
(gpu)."""
import numpy as np
import pytest

import serve
from oracle.workloads_cpu import config_oracle
import workloads as W

pytestmark = pytest.mark.gpu


def small_trace():
    # many overlapping arrivals so the batch changes membership often
    return serve.make_trace(10, rate=2000.0, seed=11, min_cells=64 * 64, max_cells=320 * 256,
                            min_steps=30, max_steps=260, bands=4)


def test_served_runs_equal_their_solo_runs():
    trace = small_trace()
    print("serve batched", [(r.w, r.h, r.steps) for r in trace], flush=True)
    s_b, res_b = serve.serve_batched(trace, every=50, block=16, keep_cells=True)
    print(s_b, flush=True)
    assert s_b["peak_resident"] > 1, "the trace never overlapped on the device"
    s_p, res_p = serve.serve_plain(trace, every=50, block=16, keep_cells=True)
    for r in trace:
        a, b = res_b[r.rid], res_p[r.rid]
        assert np.array_equal(a.final_cells, b.final_cells), r
        assert [s[0] for s in a.snapshots] == [s[0] for s in b.snapshots]
        assert len(a.snapshots) == -(-r.steps // 50)
        for sa, sb in zip(a.snapshots, b.snapshots):
            for x, y in zip(sa[1:], sb[1:]):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_served_runs_equal_their_cpu_replay(oracle):
    trace = small_trace()[:3]
    _, res = serve.serve_batched(trace, every=50, block=16, keep_cells=True)
    for r in trace:
        print(f"cpu replay {r}", flush=True)
        cfg = dict(w=r.w, h=r.h, rest=True, density=r.density, init_u=(0.0, 0.0),
                   inflow=(r.drift, 0.0), inlet=r.inlet, seed=r.seed, bands=r.bands,
                   plate=W.plate_box(r.w, r.h, r.plate_x, r.plate_rows))
        lat = config_oracle(oracle, cfg)
        # the same sampling history: Field::sample with alpha 1 every 50 steps
        # and at the end, into one running field (src/render.rs:107-115)
        f = oracle.Field(r.w, r.h, 16, 16)
        done = 0
        snaps = res[r.rid].snapshots
        for i, snap in enumerate(snaps):
            lat.advance(snap[0] - done)
            done = snap[0]
            f.sample(lat, 1.0)
            for x, y in zip(snap[1:], f.read()):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (r.rid, i)
        assert done == r.steps
        assert np.array_equal(res[r.rid].final_cells, lat.cells)
