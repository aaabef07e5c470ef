"""Host-side pieces of the serving loop: the request trace and the HBM
footprint used for admission (no device calls)."""
import math

import numpy as np

import serve


def test_trace_statistics_follow_the_served_stream_config():
    t = serve.make_trace(2000, rate=50.0, seed=3)
    gaps = np.diff([0.0] + [r.arrival for r in t])
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0          # Poisson at 50/s
    cells = np.array([r.cells for r in t], dtype=float)
    assert cells.min() >= 0.7 * 256 * 256 and cells.max() <= 1.3 * 4096 * 2304
    # log-uniform: the median sits near the geometric mean of the bounds
    gm = math.sqrt(256 * 256 * 4096 * 2304)
    assert 0.7 * gm < np.median(cells) < 1.4 * gm
    steps = np.array([r.steps for r in t])
    assert steps.min() >= 500 and steps.max() <= 20000
    for r in t:
        assert r.w % 64 == 0 and r.h % 64 == 0 and r.h % 2 == 0
        assert 0.1 <= r.drift <= 0.3 and r.w / 8 <= r.plate_x <= r.w / 3


def test_footprint_counts_planes_tables_and_field():
    r = serve.Request(0, 0.0, 4096, 2304, 1000, 1, 8, 0.2, 1000.0, 288)
    fp = serve.footprint(r, 32)
    planes = 16 * 64 * 2304 * 8          # 2 x 7 planes + solid + jump table
    assert planes < fp < planes + 64 * 2**20
    assert serve.footprint(r, 32) > serve.footprint(serve.Request(1, 0.0, 256, 256, 1000, 1, 8, 0.2, 50.0, 30), 32)
