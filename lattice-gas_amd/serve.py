"""Serving FHP runs on one MI355X with continuous batching.

A request is one lattice-gas run: size, step count, plate geometry, inflow
drift, seed and CPU band count (the band count fixes its random streams, so
each served run stays bit-exact with its own CPU replay). The server

  * admits requests as they arrive, while their HBM footprint fits the budget;
  * keeps every admitted lattice resident on the device;
  * advances all resident lattices together, one tile-table step kernel per
    step with the step loop replayed from a hipGraph (lgca_amd.Batch), in
    quanta of `every` steps (a run with fewer steps left stops early within
    the quantum);
  * streams back each run's coarse-grained velocity field every `every` of its
    own steps (blocks of `block` x `block` cells, the reference's Field with
    alpha = 1) and at its end;
  * retires finished runs, freeing their memory for the queue.

The plain loop runs the same trace one lattice at a time, first come first
served, with the same snapshots. Both report p50/p99 latency (completion minus
arrival) and aggregate throughput (cell updates over the makespan).
"""
from __future__ import annotations

import math
import queue
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from lgca_amd import Batch, GpuLattice
import workloads as W

GIB = 1 << 30


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the trace starts
    w: int
    h: int
    steps: int
    seed: int
    bands: int
    drift: float
    plate_x: float
    plate_rows: int
    density: float = 0.25
    inlet: int = 8

    @property
    def cells(self) -> int:
        return self.w * self.h


@dataclass
class Result:
    rid: int
    arrival: float
    start: float = 0.0
    done: float = 0.0
    snapshots: list = field(default_factory=list)   # (step, ux, uy, rho, solid)
    final_cells: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def make_trace(n: int, rate: float = 50.0, seed: int = 1, min_cells: int = 256 * 256,
               max_cells: int = 4096 * 2304, min_steps: int = 500, max_steps: int = 20000,
               bands: int = 8) -> list[Request]:
    """Poisson arrivals at `rate` per second; cell counts log-uniform between
    the bounds (aspect from 1:1 up to 16:9, sides multiples of 64, even
    height); steps uniform; plate 10-25% of H tall at x in [W/8, W/3];
    inflow drift uniform in [0.1, 0.3]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        cells = math.exp(rng.uniform(math.log(min_cells), math.log(max_cells)))
        aspect = rng.uniform(1.0, 16.0 / 9.0)
        w = max(64, int(round(math.sqrt(cells * aspect) / 64.0)) * 64)
        h = max(64, int(round(cells / w / 64.0)) * 64)
        steps = int(rng.integers(min_steps, max_steps + 1))
        frac = rng.uniform(0.10, 0.25)
        out.append(Request(rid=i, arrival=t, w=w, h=h, steps=steps,
                           seed=int(rng.integers(1, 2**63)), bands=bands,
                           drift=float(rng.uniform(0.1, 0.3)),
                           plate_x=float(rng.uniform(w / 8.0, w / 3.0)),
                           plate_rows=max(2, int(round(frac * h)))))
    return out


def build(r: Request, block: int) -> GpuLattice:
    g = GpuLattice(r.w, r.h, True, r.seed, r.bands)
    g.set_inlet(r.inlet, r.density, (r.drift, 0.0))
    g.add_solid_box(*W.plate_box(r.w, r.h, r.plate_x, r.plate_rows))
    g.init_equilibrium(r.density, 0.0, 0.0)
    g.attach_field(block, block)
    g.advance(0)  # builds the stream-jump tables now, not in the first quantum's call
    return g


class _Builder:
    """Builds admitted requests' lattices on a thread of its own, one at a
    time, so that the serving loop never waits for a build: a build's set-up
    calls wait for their own stream, which can share a hardware queue with
    the batch's queued quanta and then waits for those too. ctypes releases
    the interpreter lock during each call, so the loop runs meanwhile."""

    def __init__(self, block: int):
        self.block = block
        self.todo, self.built = queue.Queue(), queue.Queue()
        self.thread = threading.Thread(target=self._run, daemon=True)
        self.thread.start()

    def _run(self):
        while True:
            r = self.todo.get()
            if r is None:
                return
            try:
                self.built.put((r, build(r, self.block), None))
            except Exception as e:  # reported by the loop
                self.built.put((r, None, e))

    def submit(self, r: Request):
        self.todo.put(r)

    def take(self, wait: float = 0.0):
        """Lattices built since the last call ((request, lattice) pairs),
        waiting up to `wait` seconds for the first one."""
        out = []
        try:
            out.append(self.built.get(timeout=wait) if wait > 0 else self.built.get_nowait())
            while True:
                out.append(self.built.get_nowait())
        except queue.Empty:
            pass
        for r, g, e in out:
            if e is not None:
                raise e
        return [(r, g) for r, g, _ in out]

    def close(self):
        self.todo.put(None)
        self.thread.join()


def footprint(r: Request, block: int) -> int:
    """Device bytes of one resident run: two plane sets, the solid plane,
    jump polynomials (two per word, and the lead jump), word order, the two digit and inflow slots of the
    two-pass step, the stream tables, the field and the byte staging buffer
    of load/store."""
    wpr = (r.w + 63) // 64
    words = wpr * r.h
    iw_h = ((min(r.inlet, r.w) + 63) // 64) * r.h
    nb = -(-r.h // (-(-r.h // min(r.bands, r.h))))
    tab_steps = max(3, min(256, (32 << 20) // ((nb + 1) * 256 * 8)))
    blocks = (r.w // block) * (r.h // block)
    return (words * (8 * (14 + 1 + 2 + 1) + 4 + 80) + iw_h * (16 + 112) + 4 * r.h
            + tab_steps * (nb + 1) * 2048 + 48 * blocks + r.cells)


def _snapshot(g: GpuLattice, step: int):
    ux, uy, rho, sol = g.read_field()
    return (step, ux, uy, rho, sol)


def _percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if xs else float("nan")


def _summary(trace, results, t_first, t_last, mode):
    lat = [results[r.rid].latency for r in trace]
    work = sum(r.cells * r.steps for r in trace)
    span = t_last - t_first
    return {"mode": mode, "requests": len(trace), "cell_updates": work,
            "makespan_s": round(span, 4), "gcups": round(work / span / 1e9, 3),
            "p50_latency_s": round(_percentile(lat, 50), 4),
            "p99_latency_s": round(_percentile(lat, 99), 4),
            "mean_latency_s": round(float(np.mean(lat)), 4)}


def serve_batched(trace: list[Request], every: int = 100, block: int = 32,
                  hbm_budget: int = 200 * GIB, keep_cells: bool = False, shape=None,
                  alone_max: int = 2, depth: int = 2):
    """Continuous batching over the trace; returns (summary, results).

    Time is cut into quanta of `every` steps. At each quantum boundary the
    server admits what has arrived (while the HBM budget allows), advances
    every resident run by min(every, steps left) in the same launches, takes
    one batched snapshot of all of them, and retires the runs that are done.
    A run joins at a boundary, so its snapshots fall every `every` of its own
    steps and at its end, exactly as in a run of its own.

    Quanta are queued `depth` deep (Batch.advance_sampling_async): the next
    quantum is queued -- on a new batch when the membership changes, whose
    queue waits for the last one on the device -- before the host waits for
    the oldest one's snapshot, so the device has the next quantum while the
    host reads snapshots and retires runs. Arrivals are admitted while the
    HBM budget allows and built on a thread of their own (_Builder); a built
    run joins the next quantum queued. A run's steps count when they are
    queued; it is retired when its last quantum is read.
    Finished runs' lattices are destroyed at once (their memory goes back to
    the library's cache; a batch still holding one frees it with the batch),
    and their bytes leave the budget.

    With at most `alone_max` runs on the device and nothing queued (the start
    and the tail of a stream), runs whose lattices take resident bands
    advance on their own launches instead of in a batch: for so few lattices
    those beat the batch kernel (4096 x 2304: ~8.5 us per step resident
    against ~17 us in a batch of one)."""
    pending = deque(sorted(trace, key=lambda r: r.arrival))
    results = {r.rid: Result(r.rid, r.arrival) for r in trace}
    runs = []               # dicts: req, g, sched (steps queued), done (steps read back), bytes
    used = 0                # bytes of the runs' lattices
    inflight = deque()      # queued quanta: (batch, runs, limits)
    cur, cur_ids = None, None   # batch of the latest queued quantum and its members
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    quanta = 0
    peak = 0
    spent = {"admit": 0.0, "batch_setup": 0.0, "advance": 0.0, "snapshot": 0.0,
             "retire": 0.0, "idle": 0.0}  # wall seconds by activity

    builder = _Builder(block)
    building = 0            # admitted requests not yet built

    def admit_arrived(wait=0.0):
        nonlocal used, building
        t = time.perf_counter()
        while pending and pending[0].arrival <= clock():
            r = pending[0]
            fp = footprint(r, block)
            if used + fp > hbm_budget and (runs or building):
                break                                # wait for memory to free up
            pending.popleft()
            used += fp
            results[r.rid].start = clock()
            builder.submit(r)
            building += 1
        for r, g in builder.take(wait if building else 0.0):
            building -= 1
            runs.append({"req": r, "g": g, "sched": 0, "done": 0, "bytes": footprint(r, block)})
        spent["admit"] += time.perf_counter() - t

    def close_batch(b):
        if b is not None and b is not cur and all(q[0] is not b for q in inflight):
            b.close()

    while pending or runs or inflight or building:
        admit_arrived()
        nxt = [a for a in runs if a["sched"] < a["req"].steps]
        peak = max(peak, len(runs))
        if nxt and len(inflight) < depth:
            limits = [min(every, a["req"].steps - a["sched"]) for a in nxt]
            if (not inflight and len(nxt) <= alone_max
                    and all(a["g"].resident() is not None for a in nxt)):
                # a few runs left (the start and the tail of a stream): each
                # one's own resident-band launches beat the batch kernel over
                # so few lattices; same steps, same snapshot (alpha 1: the
                # field is the coarse-grain of the quantum's last state)
                old, cur, cur_ids = cur, None, None
                close_batch(old)
                ta = time.perf_counter()
                for a, lim in zip(nxt, limits):
                    a["g"].advance_async(lim, every, 1.0)
                for a in nxt:
                    a["g"].wait()
                spent["advance"] += time.perf_counter() - ta
                quanta += 1
                ts = time.perf_counter()
                snaps = [a["g"].read_field() for a in nxt]
                spent["snapshot"] += time.perf_counter() - ts
                for a, lim in zip(nxt, limits):
                    a["sched"] += lim
                inflight.append((None, nxt, limits, snaps))
            else:
                ids = [a["req"].rid for a in nxt]
                if ids != cur_ids:
                    ts = time.perf_counter()
                    old, cur = cur, Batch([a["g"] for a in nxt])
                    cur_ids = ids
                    if shape:
                        cur.set_shape(*shape)
                    close_batch(old)
                    spent["batch_setup"] += time.perf_counter() - ts
                ta = time.perf_counter()
                cur.advance_sampling_async(every, limits, 1.0)
                spent["advance"] += time.perf_counter() - ta
                quanta += 1
                for a, lim in zip(nxt, limits):
                    a["sched"] += lim
                inflight.append((cur, nxt, limits, None))
                if len(inflight) < depth:
                    continue                         # queue the next quantum before reading
        if not inflight:
            if building:                             # nothing to step until a build is done
                admit_arrived(0.002)
            elif pending:
                t = time.perf_counter()
                time.sleep(max(0.0, min(pending[0].arrival - clock(), 0.01)))
                spent["idle"] += time.perf_counter() - t
            continue
        b, members, limits, snaps = inflight.popleft()
        if snaps is None:
            ts = time.perf_counter()
            snaps = b.read_snapshot()                # waits for this quantum only
            spent["snapshot"] += time.perf_counter() - ts
        tr = time.perf_counter()
        for a, lim, snap in zip(members, limits, snaps):
            a["done"] += lim
            results[a["req"].rid].snapshots.append((a["done"],) + tuple(snap))
        finished = [a for a in members if a["done"] >= a["req"].steps]
        for a in finished:
            res = results[a["req"].rid]
            if keep_cells:
                res.final_cells = a["g"].store()
            res.done = clock()
            a["g"].close()                           # a batch that holds it frees it later
            used -= a["bytes"]
        if finished:
            gone = {id(a) for a in finished}
            runs = [a for a in runs if id(a) not in gone]
        close_batch(b)
        spent["retire"] += time.perf_counter() - tr
    builder.close()
    if cur is not None:
        cur.close()
    t_last = max(res.done for res in results.values())
    s = _summary(trace, results, 0.0, t_last, "batched")
    s["quanta"] = quanta
    s["peak_resident"] = peak
    s["queue_depth"] = depth
    s["wall_s_by_activity"] = {k: round(v, 4) for k, v in spent.items()}
    return s, results


def serve_plain(trace: list[Request], every: int = 100, block: int = 32, keep_cells: bool = False):
    """The plain loop: one lattice at a time, first come first served."""
    results = {r.rid: Result(r.rid, r.arrival) for r in trace}
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    spent = {"admit": 0.0, "advance": 0.0, "snapshot": 0.0, "idle": 0.0}
    for r in sorted(trace, key=lambda r: r.arrival):
        wait = r.arrival - clock()
        if wait > 0:
            time.sleep(wait)
            spent["idle"] += wait
        res = results[r.rid]
        res.start = clock()
        t = time.perf_counter()
        g = build(r, block)
        spent["admit"] += time.perf_counter() - t
        done = 0
        while done < r.steps:
            q = min(every, r.steps - done)
            t = time.perf_counter()
            # the quantum and its sample in one call, as the reference's run
            # loop drives GpuLattice::advance_sampling (src/sim.rs:119-143);
            # resident bands take the sample in the same launch
            g.advance_sampling(q, every, 1.0)
            spent["advance"] += time.perf_counter() - t
            done += q
            t = time.perf_counter()
            res.snapshots.append(_snapshot(g, done))
            spent["snapshot"] += time.perf_counter() - t
        if keep_cells:
            res.final_cells = g.store()
        res.done = clock()
        g.close()
    t_last = max(res.done for res in results.values())
    s = _summary(trace, results, 0.0, t_last, "plain")
    s["wall_s_by_activity"] = {k: round(v, 4) for k, v in spent.items()}
    return s, results
