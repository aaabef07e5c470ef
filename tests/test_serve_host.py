"""Host-side pieces of the serving loop: the request trace and the HBM
footprint used for admission (no device calls)."""
import math
import threading
import time

import numpy as np
import pytest

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


class _Lattice:
    """Stand-in for a resident lattice: counts its own steps. With `bands`
    set it says it runs resident bands, so that a run with few others on the
    device advances on its own (advance_async / wait / read_field). `last`
    is the batch advance that stepped it last (a thread), which the next one
    waits for, as a batch's queue waits on the device for the members'
    earlier work."""
    live = {}
    bands = False
    alone_calls = 0

    def __init__(self, r):
        self.r, self.steps, self.closed, self.queued, self.last = r, 0, False, 0, None
        _Lattice.live[r.rid] = self

    def resident(self):
        return (1, 1) if _Lattice.bands else None

    def advance_async(self, n, every, alpha):
        assert not self.closed and 0 < n <= every and alpha == 1.0
        assert self.last is None or not self.last.is_alive(), "own call before the batch's steps"
        self.queued = n
        _Lattice.alone_calls += 1

    def wait(self):
        time.sleep(0.001)               # the quantum's steps on the device
        self.steps += self.queued
        self.queued = 0

    def read_field(self):
        return (np.full(1, self.steps, np.float32),) * 4

    def close(self):
        assert not self.closed, "a lattice freed twice"
        self.closed = True
        del _Lattice.live[self.r.rid]


class _Batch:
    """Stand-in for lgca_amd.Batch: each sampling advance runs on a thread of
    its own, which first waits for the advances that stepped its members last
    (queued on this or another batch) and sleeps, so that the host work the
    server overlaps with queued quanta really runs beside them; at most two
    snapshots wait to be read, oldest first."""

    def __init__(self, members):
        assert all(not m.closed for m in members)
        self.members, self.closed, self.queue, self.errors = members, False, [], []

    def _advance(self, before, n, limits, out):
        for t in before:
            t.join()
        time.sleep(0.002)
        for m, lim in zip(self.members, limits):
            if m.closed:
                self.errors.append("a member freed before its steps")
            m.steps += min(n, lim)
        out.extend((np.full(1, m.steps, np.float32),) * 4 for m in self.members)

    def advance_sampling_async(self, n, limits, alpha):
        assert not self.closed and len(limits) == len(self.members)
        assert len(self.queue) < 2, "a third unread snapshot"
        before = {id(m.last): m.last for m in self.members if m.last is not None}.values()
        out = []
        t = threading.Thread(target=self._advance, args=(list(before), n, list(limits), out))
        for m in self.members:
            m.last = t
        t.start()
        self.queue.append((t, out))

    def read_snapshot(self):
        t, out = self.queue.pop(0)
        t.join()
        assert not self.errors, self.errors
        return out

    def close(self):
        assert not self.closed
        for t, _ in self.queue:
            t.join()
        assert not self.errors, self.errors
        self.closed = True


@pytest.mark.parametrize("bands", [False, True])
def test_batched_scheduler_snapshots_budget_and_frees(monkeypatch, bands):
    """The continuous-batching loop with stand-ins for the device objects,
    quanta queued two deep: every run's snapshots fall at its own multiples of
    `every` and its last step, every lattice is freed exactly once and never
    before its last steps, and the bytes of lattices not yet freed stay
    within the budget whenever a run is admitted beside others."""
    rng = np.random.default_rng(5)
    trace = [serve.Request(i, 0.002 * i, 64, 64, int(rng.integers(1, 9)) * 25 + int(rng.integers(0, 25)),
                           i + 1, 4, 0.2, 10.0, 4) for i in range(40)]
    size = {r.rid: int(rng.integers(1, 5)) for r in trace}
    budget = 6
    high = [0]

    def build(r, block):
        _Lattice(r)
        live = sum(size[k] for k in list(_Lattice.live))
        if len(_Lattice.live) > 1:
            high[0] = max(high[0], live)
        return _Lattice.live[r.rid]

    monkeypatch.setattr(_Lattice, "bands", bands)
    monkeypatch.setattr(_Lattice, "alone_calls", 0)
    monkeypatch.setattr(serve, "build", build)
    monkeypatch.setattr(serve, "Batch", _Batch)
    monkeypatch.setattr(serve, "footprint", lambda r, block: size[r.rid])
    s, res = serve.serve_batched(trace, every=25, block=16, hbm_budget=budget)
    assert not _Lattice.live
    assert 0 < high[0] <= budget
    for r in trace:
        steps = [snap[0] for snap in res[r.rid].snapshots]
        assert steps == list(range(25, r.steps, 25)) + [r.steps], r
        assert all(int(snap[1][0]) == snap[0] for snap in res[r.rid].snapshots)
        assert res[r.rid].done >= res[r.rid].start >= 0.0
    assert s["requests"] == len(trace) and s["peak_resident"] >= 2
    # with resident bands, runs with at most one other on the device advance
    # on their own (the start and the tail of the stream here); without, never
    assert (_Lattice.alone_calls > 0) == bands
