#!/usr/bin/env python3
"""Throughput of the FHP step on one MI355X, with its HBM roofline and the
reference CPU path timed on the same host.

Workload (one GPU): the 9.4M-cell default-scale run, 4096 x 2304 sites with
rest particles, links and rest Bernoulli(0.25) from seed 42, a plate 288 rows
tall at x = 1024, 8 inflow columns re-seeded each step at drift 0.2, and the
32 x 32 coarse-grain (exact integer block sums + f32 field) every 100 steps.
A bench "step" is one lattice step. The lattice is built and resident in HBM
before the timed region starts.

Prints one JSON line. `value` is cell updates per second (GCUPS) over the
timed steps; `roofline.achieved` is the algorithmic traffic of one step
(15 bits per cell: 7 planes read, 7 written, the solid plane read) divided by
the mean device time per step, from HIP event pairs on the engine's stream
around each launch of steps (at this size the resident-band kernel: one
cooperative launch per chunk of steps, a chunk ending at a sample, which the
launch itself takes); `roofline.traffic` is the HBM bytes of one such launch
from the committed rocprofv3 summary of the same kernel and workload.
`cpu_baseline` is the CPU oracle (a restatement of the reference's CPU path,
one thread per row band, spawned per step as the reference does) replaying
the same run on this host; its final cells are compared with the GPU's.

Further objects in the same line: `small_512x256` (the reference's 512 x 256
CPU-runnable case on the GPU; its CPU rates and check sit under
`cpu_baseline`), `shear_537m` (the 32768 x 16384 transport workload's step
rate), `ensemble` (256 uneven lattices, batched vs one at a time),
`serving` (256 Poisson requests offered below the batched server's capacity,
continuous batching vs the plain loop, with the offered load and whether each
server keeps up), `serving_saturated` (64 requests at 50/s, above capacity: the
servers' throughput when the queue never empties),
`transport_measure` (the reference's --measure run on the engine: viscosity and
advection factor with their wall time), `frames` (the frame writers on the
reference's render-benchmark field, and what queueing a frame every 500 steps
to the writer pool costs the headline run's stepping loop) and
`reference_suite` (the reference's own GPU benchmark cases, same lattices and
harness, beside its published figures) and `autotune` (the pipeline lg_autotune
keeps for the headline workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lattice-gas_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
BYTES_PER_CELL_UPDATE = 15 / 8  # 7 + 7 + 1 bits
METRIC = "GCUPS at 9.4M cells (% HBM roofline); served-batch GCUPS + p99 vs plain loop"


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))  # the box's CPU share is 16 cores


def cpu_baseline(cfg, total_steps, gpu_cells, budget_s):
    """Replay the GPU run on the oracle with one thread per band; time per
    step over the replay (best of the 7 chunks, as benches/speed.rs:160-179
    reports best-of), and check the final cells. Bounded by `budget_s`:
    a replay that would exceed it is cut, and the check skipped."""
    import numpy as np

    from oracle import oracle as O
    from oracle.workloads_cpu import config_oracle

    lat = config_oracle(O, cfg)
    lat.set_parallel(True)
    cells = cfg["w"] * cfg["h"]
    t0 = time.perf_counter()
    lat.advance(2)  # warm-up
    per_step = (time.perf_counter() - t0) / 2
    replay = (total_steps - 2) * per_step <= budget_s
    n_left = (total_steps - 2) if replay else max(7, int(budget_s / max(per_step, 1e-6)))
    rates = []
    done = 0
    for i in range(7):                  # seven chunks of (nearly) equal length
        k = n_left * (i + 1) // 7 - done
        if k == 0:
            continue
        t = time.perf_counter()
        lat.advance(k)
        rates.append(cells * k / (time.perf_counter() - t))
        done += k
    verified = None
    if replay and gpu_cells is not None:
        verified = bool(np.array_equal(lat.cells, gpu_cells))
    rates.sort()
    best = rates[-1]
    # one thread (the same bands run one after another), best of 7 short chunks
    lat.set_parallel(False)
    one = []
    k1 = max(1, int(min(budget_s / 6.0, 5.0) / max(per_step * cfg["bands"], 1e-6) / 7))
    for _ in range(7):
        t = time.perf_counter()
        lat.advance(k1)
        one.append(cells * k1 / (time.perf_counter() - t))
    return {
        "value": round(best / 1e9, 4),
        "unit": "GCUPS",
        "cores": cfg["bands"],
        "kind": "port",
        "sample": (f"oracle replay of the same {cfg['w']}x{cfg['h']} run, "
                   f"{done + 2} steps on {cfg['bands']} threads (one per band), "
                   f"best of {len(rates)} chunks; median {rates[len(rates) // 2] / 1e9:.4f} GCUPS"),
        "verified_bit_exact": verified,
        "single_thread": {"value": round(max(one) / 1e9, 4), "unit": "GCUPS", "cores": 1,
                          "sample": f"same run continued on one thread, best of 7 x {k1} steps"},
    }


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, as recorded in a
    profile summary under profiles/ by a line
        traffic <kernel> <workload> bytes_per_launch=<N>
    (rocprofv3 FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    same counter's reading for a plain copy of known size); None if absent."""
    import glob
    import re

    pat = re.compile(r"^traffic (\S+) (\S+) bytes_per_launch=(\d+)\s*$")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.match(line.strip())
                if m and m.group(1) == kernel and m.group(2) == workload:
                    return int(m.group(3))
    return None


def ensemble_leg(n, steps, bands, seed=4):
    """Batched ensemble: n independent lattices, widths and heights uniform on
    256..2048 in steps of 64, per-lattice seed, plate 10-25% of H tall at x in
    [W/8, W/3], inflow drift in [0.1, 0.3], `steps` steps each; one tile-table
    step kernel per step over all of them, replayed from a hipGraph with an
    autotuned shape, against the plain loop that runs every one of them, one
    at a time, for the same steps."""
    from lgca_amd import Batch
    import workloads as W

    specs = W.config4(n, bands=bands, seed=seed, steps=steps)
    build = W.config_gpu

    cells = sum(sp["w"] * sp["h"] for sp in specs)
    lats = [build(sp) for sp in specs]
    b = Batch(lats)
    t0 = time.perf_counter()
    used = b.autotune(budget_steps=min(512, steps // 4))
    b.advance(steps - used)
    t_batch = time.perf_counter() - t0
    shape = b.shape()
    passes, units, dunits = b.pipeline()
    waves = b.waves()
    b.close()
    # the plain loop over the same lattices (rebuilt), one at a time
    t_plain = 0.0
    plain_cells = 0
    probe = n
    for sp in specs:
        g = build(sp)
        t = time.perf_counter()
        g.advance(steps)
        t_plain += time.perf_counter() - t
        plain_cells += sp["w"] * sp["h"]
        g.close()
    gc_b = cells * steps / t_batch / 1e9
    gc_p = plain_cells * steps / t_plain / 1e9
    return {"lattices": n, "steps": steps, "cells": cells, "bands": bands,
            "batched_gcups": round(gc_b, 3), "plain_loop_gcups": round(gc_p, 3),
            "speedup": round(gc_b / gc_p, 3), "batched_s": round(t_batch, 4),
            "autotune_steps": used,
            "autotuned_shape": {"tile": shape[0], "grid": shape[1], "chunk": shape[2],
                                "passes": passes, "cells_per_lane": units,
                                "draw_cells_per_lane": dunits, "waves_per_simd": waves},
            "plain_loop_sample": f"all {probe} lattices, {steps} steps each"}


def autotune_leg(bands):
    """lg_autotune on a lattice of the headline workload (built apart from the
    timed one): which step pipeline -- the size rule, resident bands on or
    off, one-launch two-pass, draw-ahead, fused, temporally blocked collides
    of depth 2, 4, 8 -- it keeps, and the rate of the kept pipeline over 200
    more steps."""
    import workloads as W

    cfg = W.config2(bands=bands)
    g = W.config_gpu(cfg)
    t0 = time.perf_counter()
    used = g.autotune(0)
    t_tune = time.perf_counter() - t0
    passes, cpl, dcpl, ahead = g.pipeline()
    depth, rows = g.temporal_depth()
    res = g.resident()
    g.advance(20)
    t0 = time.perf_counter()
    g.advance(200)
    dt = time.perf_counter() - t0
    g.close()
    return {"workload": cfg["name"], "used_steps": used, "tune_s": round(t_tune, 4),
            "kept": {"resident_bands": list(res) if res else None, "passes": passes,
                     "draw_ahead": ahead, "temporal_depth": depth},
            "kept_gcups_200_steps": round(cfg["w"] * cfg["h"] * 200 / dt / 1e9, 3)}


class _Skip(Exception):
    pass


def small_leg(bands, passes, units, dunits, ahead):
    """The reference's CPU-runnable case (512 x 256, plate 64 rows at x = 128,
    2000 steps, 16 x 16 block sums) on the GPU: its rate, and its final cells
    and exact block sums for the CPU baseline to check."""
    import workloads as W

    cfg = W.config1(bands=bands)
    cells = cfg["w"] * cfg["h"]
    g = W.config_gpu(cfg)
    g.set_work_unit(units)
    g.set_passes(passes)
    g.set_draw_unit(dunits)
    if ahead >= 0:
        g.set_draw_ahead(ahead)
    g.attach_field(cfg["block"], cfg["block"])
    t0 = time.perf_counter()
    g.advance_sampling(cfg["steps"], cfg["steps"], 1.0)
    wall = time.perf_counter() - t0
    final = (g.store(), g.read_block_sums())
    g.close()
    return ({"workload": cfg["name"], "cells": cells, "steps": cfg["steps"],
             "gpu_gcups": round(cells * cfg["steps"] / wall / 1e9, 3)}, cfg, final)


def cpu_small(cfg, final):
    """The 512 x 256 case on the oracle, all host cores (one thread per band)
    and one thread, each over the full 2000 steps; the GPU's final cells and
    block sums compared with the oracle's."""
    import numpy as np

    from oracle import oracle as O
    from oracle.workloads_cpu import config_oracle

    cells = cfg["w"] * cfg["h"]
    rates = {}
    for par in (True, False):
        lat = config_oracle(O, cfg)
        lat.set_parallel(par)
        t0 = time.perf_counter()
        lat.advance(cfg["steps"])
        rates[par] = cells * cfg["steps"] / (time.perf_counter() - t0) / 1e9
    f = O.Field(cfg["w"], cfg["h"], cfg["block"], cfg["block"])
    same = bool(np.array_equal(final[0], lat.cells)) and all(
        np.array_equal(a, b) for a, b in zip(final[1], f.block_sums(lat)))
    return {"workload": cfg["name"], "value": round(rates[True], 4), "unit": "GCUPS",
            "cores": cfg["bands"], "single_thread": round(rates[False], 4),
            "verified_bit_exact": same}


def shear_leg(steps, bands, passes, units, dunits, ahead):
    """The transport-measure workload at full size (32768 x 16384, 537M cells,
    periodic, density 0.3, shear wave from seed 7): step rate and algorithmic
    HBM rate of the step kernels, with the lattice (940 MB of planes, beyond
    the 256 MB MALL) resident before the timed region."""
    import workloads as W

    cfg = W.config3(bands=bands)
    g = W.config_gpu(cfg)
    g.set_work_unit(units)
    g.set_passes(passes)
    g.set_draw_unit(dunits)
    if ahead >= 0:
        g.set_draw_ahead(ahead)
    g.advance(2)  # warm-up (builds the tables, first launches)
    t0 = time.perf_counter()
    dev_ms = g.advance_timed(steps)
    wall = time.perf_counter() - t0
    cells = cfg["w"] * cfg["h"]
    p = g.pipeline()
    pipe = {"passes": p[0], "cells_per_lane": p[1],
            "draw_cells_per_lane": p[2] if p[0] == 2 else None}
    g.close()
    step_ms = dev_ms / steps
    ach = cells * BYTES_PER_CELL_UPDATE / (step_ms * 1e-3) / 1e9
    return {"workload": cfg["name"], "cells": cells, "steps": steps, "pipeline": pipe,
            "gcups": round(cells * steps / wall / 1e9, 3), "mean_step_ms": round(step_ms, 4),
            "achieved_gbs": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4)}


def transport_leg():
    """The reference's --measure (src/main.rs:527-548 → src/transport.rs:291-319)
    on the engine: viscosity and advection factor of the 256 x 256 boxes at
    density 0.22 with rest particles, CPU threads 12, seed 0x5EED, whose
    published values are 0.2989 and 0.4405 (docs/transport.md:14-15)."""
    import lgca_amd

    t0 = time.perf_counter()
    nu, g = lgca_amd.measure_transport(0.22, True, 12, 0x5EED, 256)
    wall = time.perf_counter() - t0
    # the same procedure at the shear workload's statistics (density 0.3 with
    # rest particles, 16 bands, seed 7)
    t0 = time.perf_counter()
    nu3, g3 = lgca_amd.measure_transport(0.3, True, 16, 7, 256)
    wall3 = time.perf_counter() - t0
    return {"nu": round(nu, 6), "g": round(g, 6), "wall_s": round(wall, 3),
            "published": {"nu": 0.2989, "g": 0.4405},
            "matches_published": f"{nu:.4f}" == "0.2989" and f"{g:.4f}" == "0.4405",
            "shear_workload_statistics": {"density": 0.3, "bands": 16, "seed": 7,
                                          "nu": round(nu3, 6), "g": round(g3, 6),
                                          "wall_s": round(wall3, 3)}}


def frames_leg(bands, passes, units, dunits, ahead, workers=4, frames=10, frame_every=500):
    """Frame output (src/main.rs:196-281), measured twice.

    Writers: the field the reference times its encoders on (2048 x 1280 plate
    run, 50 steps, blocks of 20), written by the library's write_vorticity
    (scale 8, wmax 0.01) and write_arrows (1.6 in the frame (0.4, 0)), best
    and median of 15 writes each, beside the reference's published 6.490 ms
    and 2.988 ms (docs/benchmarks.md:312-313, an M3 Pro).

    Run loop: the headline 4096 x 2304 workload stepped `frames` x
    `frame_every` steps (coarse-grain every 100 steps) with and without a
    frame queued after each chunk to `workers` writer threads (blocks of 32,
    scale 8): what writing frames costs the stepping loop. Returns the
    measures and the writers' field, for the CPU baseline's encoders."""
    import shutil
    import tempfile

    import lgca_amd
    import workloads as W

    rc = W.render_bench(bands)
    g = W.config_gpu(rc)
    g.attach_field(rc["block"], rc["block"])
    g.advance(rc["steps"])
    g.sample_field(1.0)
    ux, uy, _, solid = g.read_field()
    bw, bh, bx, by = g.field_dims()
    g.close()
    field = (ux, uy, solid, bw, bh, bx, by)
    d = tempfile.mkdtemp(prefix="lgca-frames-")
    try:
        t_png, t_svg = [], []
        for _ in range(15):
            t = time.perf_counter()
            lgca_amd.write_vorticity(os.path.join(d, "v.png"), field, 8, 0.01)
            t_png.append(time.perf_counter() - t)
            t = time.perf_counter()
            lgca_amd.write_arrows(os.path.join(d, "a.svg"), field, 1.6, (0.4, 0.0))
            t_svg.append(time.perf_counter() - t)
        t_png.sort()
        t_svg.sort()
        py = max(int(round(8 * 0.8660254 * by / bx)), 1)
        pixels = bw * 8 * bh * py
        writers = {"workload": rc["name"], "field": f"{bw}x{bh} blocks of {bx}",
                   "image": f"{bw * 8}x{bh * py}",
                   "png_ms": round(t_png[0] * 1e3, 3), "png_median_ms": round(t_png[7] * 1e3, 3),
                   "png_mpx_per_s": round(pixels / t_png[0] / 1e6, 1),
                   "svg_ms": round(t_svg[0] * 1e3, 3), "svg_median_ms": round(t_svg[7] * 1e3, 3),
                   "reference_published_m3pro": {"png_ms": 6.490, "svg_ms": 2.988}}

        cfg = W.config2(bands=bands)
        every = cfg["sample_every"]
        g = W.config_gpu(cfg)
        g.set_work_unit(units)
        g.set_passes(passes)
        g.set_draw_unit(dunits)
        if ahead >= 0:
            g.set_draw_ahead(ahead)
        g.attach_field(cfg["block"], cfg["block"])
        g.advance_sampling(every, every, 1.0)  # warm-up
        t0 = time.perf_counter()
        for _ in range(frames):
            g.advance_sampling(frame_every, every, 1.0)
        bare = time.perf_counter() - t0
        fr = lgca_amd.Frames(workers)
        t0 = time.perf_counter()
        for k in range(frames):
            g.advance_sampling(frame_every, every, 1.0)
            fr.write_lattice(g, os.path.join(d, f"vorticity-{k:04}.png"),
                             os.path.join(d, f"arrows-{k:04}.svg"), 8, 1.6)
        queued = time.perf_counter() - t0
        fr.finish()
        on_disk = time.perf_counter() - t0
        bwf, bhf, bxf, byf = g.field_dims()
        g.close()
        pyf = max(int(round(8 * 0.8660254 * byf / bxf)), 1)
        loop = {"workload": cfg["name"], "frames": frames, "frame_every": frame_every,
                "writers": workers, "image": f"{bwf * 8}x{bhf * pyf}",
                "steps_only_s": round(bare, 4), "with_frames_s": round(queued, 4),
                "frames_on_disk_s": round(on_disk, 4),
                "loop_overhead": round(queued / bare - 1.0, 4)}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"writers": writers, "run_loop": loop}, field


def cpu_frames(field):
    """The oracle's encoders (a plain restatement of src/render.rs and
    src/png.rs, one thread) on the writers' field, best of 5 each, and the
    library's files compared with them byte for byte."""
    import shutil
    import tempfile

    import lgca_amd
    from oracle import oracle as O

    ux, uy, solid, bw, bh, bx, by = field
    t_png, t_svg = [], []
    for _ in range(5):
        t = time.perf_counter()
        png = O.vorticity_png(ux, uy, solid, bw, bh, bx, by, 8, 0.01)
        t_png.append(time.perf_counter() - t)
        t = time.perf_counter()
        svg = O.arrows_svg(ux, uy, solid, bw, bh, bx, by, 1.6, (0.4, 0.0))
        t_svg.append(time.perf_counter() - t)
    d = tempfile.mkdtemp(prefix="lgca-frames-")
    try:
        lgca_amd.write_vorticity(os.path.join(d, "v.png"), field, 8, 0.01)
        lgca_amd.write_arrows(os.path.join(d, "a.svg"), field, 1.6, (0.4, 0.0))
        with open(os.path.join(d, "v.png"), "rb") as f1, open(os.path.join(d, "a.svg"), "rb") as f2:
            same = f1.read() == png and f2.read() == svg
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"png_ms": round(min(t_png) * 1e3, 3), "svg_ms": round(min(t_svg) * 1e3, 3),
            "cores": 1, "kind": "port", "verified_bit_exact": same}


# benches/speed.rs:477-591 and their published M3 Pro (Metal) figures,
# docs/benchmarks.md:298-306: (case, time per iteration in s, Mcell/s).
REFERENCE_GPU_SUITE = [
    ("step/gpu/2048x1280/one-per-submit", 183.83e-6, 14260.5),
    ("step/gpu/2048x1280/batched", 3.434e-3, 76339.7),
    ("step/gpu/256x256/batched", 834.60e-6, 7852.4),
    ("step/gpu/4096x2048/batched", 11.109e-3, 75511.8),
    ("step/gpu/2048x1280/batched+inlet", 3.591e-3, 73009.7),
    ("step/gpu/2048x1280/batched+inlet+sample", 3.770e-3, 69532.0),
    ("field/gpu/sample/2048x1280", 150.69e-6, 17396.3),
    ("lattice/gpu/store/2048x1280", 998.14e-6, 2626.3),
    ("lattice/gpu/total-particles/2048x1280", 165.75e-6, 15815.9),
]


def _speed_case(body, warm_s=0.12, sample_s=0.15, samples=7):
    """The harness of benches/speed.rs:128-185: a warm-up, then `samples`
    samples of about `sample_s` each; the fastest sample's time per
    iteration, and the median beside it."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < warm_s or n == 0:
        body()
        n += 1
    per = (time.perf_counter() - t0) / n
    iters = max(1, int(sample_s / max(per, 1e-9)))
    res = []
    for _ in range(samples):
        t = time.perf_counter()
        for _ in range(iters):
            body()
        res.append((time.perf_counter() - t) / iters)
    res.sort()
    return res[0], res[len(res) // 2]


def reference_suite_leg(bands):
    """The reference's own GPU benchmark cases (benches/speed.rs:477-591) on
    the engine, with the same lattices (rest particles, equilibrium at
    density 0.22 and drift 0.4 from seed 0x5EED, periodic unless the case
    adds the 8-column inflow; 20 x 20 field blocks, a sample every 5 steps at
    alpha 0.2; 100 steps per timed call for the batched cases) and the same
    harness, beside the figures the reference publishes for its Metal path on
    an M3 Pro. The engine draws the CPU path's exact random streams; the
    reference's Metal kernels draw cheaper ones of their own."""
    import numpy as np

    from lgca_amd import GpuLattice

    def lattice(w, h, inlet=False, field=False):
        g = GpuLattice(w, h, True, 0x5EED, bands)
        g.init_equilibrium(0.22, 0.4, 0.0)
        if inlet:
            g.set_inlet(8, 0.22, (0.4, 0.0))
        if field:
            g.attach_field(20, 20)
        return g

    out = []
    open_handles = []

    def case(name, g, body, cells):
        best, med = _speed_case(body)
        _, ref_t, ref_rate = next(c for c in REFERENCE_GPU_SUITE if c[0] == name)
        rate = cells / best / 1e6
        out.append({"case": name, "time_s": best, "median_s": med,
                    "mcells_per_s": round(rate, 1),
                    "reference_m3pro": {"time_s": ref_t, "mcells_per_s": ref_rate},
                    "speedup": round(rate / ref_rate, 2)})
        g.close()

    bw, bh = 2048, 1280
    g = lattice(bw, bh)
    case("step/gpu/2048x1280/one-per-submit", g, lambda: g.advance(1), bw * bh)
    g = lattice(bw, bh)
    case("step/gpu/2048x1280/batched", g, lambda: g.advance(100), bw * bh * 100)
    g = lattice(256, 256)
    case("step/gpu/256x256/batched", g, lambda: g.advance(100), 256 * 256 * 100)
    g = lattice(4096, 2048)
    case("step/gpu/4096x2048/batched", g, lambda: g.advance(100), 4096 * 2048 * 100)
    g = lattice(bw, bh, inlet=True)
    case("step/gpu/2048x1280/batched+inlet", g, lambda: g.advance(100), bw * bh * 100)
    g = lattice(bw, bh, inlet=True, field=True)
    case("step/gpu/2048x1280/batched+inlet+sample", g,
         lambda: g.advance_sampling(100, 5, 0.2), bw * bh * 100)
    g = lattice(bw, bh, field=True)
    case("field/gpu/sample/2048x1280", g, lambda: g.sample_field(0.2), bw * bh)
    g = lattice(bw, bh)
    buf = np.zeros(bw * bh, dtype=np.uint8)
    case("lattice/gpu/store/2048x1280", g, lambda: g.store(buf), bw * bh)
    g = lattice(bw, bh)
    case("lattice/gpu/total-particles/2048x1280", g, g.total_particles, bw * bh)
    return {"bands": bands, "cases": out}


def serving_leg(n, rate, seed=5):
    """Served stream: Poisson arrivals, log-uniform sizes 256x256..4096x2304,
    steps 500-20000, random plate and drift, HBM budget 200 GB, 32x32-block
    velocity snapshot every 100 steps; continuous batching against the plain
    sequential loop on the same trace.

    `offered_gcups` is the trace's cell updates over the span of its arrivals:
    the rate the stream asks for. `drain_after_last_arrival_s` is how long a
    server runs on after the last request has arrived: a server that keeps up
    has only the last few requests' own run time left then; an overloaded one
    has a backlog that grew for as long as requests kept arriving, and its
    latencies measure the trace's length, not a steady state. `state` says
    which: "keeps up" when the drain is under a tenth of the arrival span."""
    import serve

    trace = serve.make_trace(n, rate=rate, seed=seed)
    span = trace[-1].arrival - trace[0].arrival
    offered = sum(r.cells * r.steps for r in trace) / span / 1e9 if span > 0 else float("inf")
    s_b, _ = serve.serve_batched(trace, every=100, block=32)
    s_p, _ = serve.serve_plain(trace, every=100, block=32)
    for s in (s_b, s_p):
        drain = s["makespan_s"] - trace[-1].arrival
        s["drain_after_last_arrival_s"] = round(drain, 4)
        s["state"] = "keeps up" if drain < 0.1 * trace[-1].arrival else "overloaded"
    return {"requests": n, "rate_per_s": rate, "arrival_span_s": round(span, 3),
            "offered_gcups": round(offered, 1), "batched": s_b, "plain_loop": s_p,
            "throughput_ratio": round(s_b["gcups"] / s_p["gcups"], 3)}


def valu_probe_table():
    """Issue rates on this GPU (lg_probe_valu): wave-instructions per CU per
    shader clock for each instruction, 8 independent chains per lane and one
    dependent chain, at 1-8 waves per SIMD; cells per CU per clock of the
    draw work's exact stream. Printed as a markdown table, one line per
    probe, flushed as it goes."""
    import lgca_amd

    wps_ind = (1, 2, 4, 8)
    print("| instruction | " + " | ".join(f"8 chains, {w} w/SIMD" for w in wps_ind)
          + " | 1 chain, 1 w/SIMD | 1 chain, 4 w/SIMD | clock GHz |", flush=True)
    print("|---" * (len(wps_ind) + 4) + "|", flush=True)
    for op in range(lgca_amd.PROBE_OPS):
        row, name, ghz = [], "", 0.0
        for w in wps_ind:
            r, ghz, name = lgca_amd.probe_valu(op, w, 8, 2048)
            row.append(r)
        for w in (1, 4):
            r, _, _ = lgca_amd.probe_valu(op, w, 1, 4096)
            row.append(r)
        print(f"| {name} | " + " | ".join(f"{r:.3f}" for r in row) + f" | {ghz:.3f} |", flush=True)
    wps = (1, 2, 3, 4, 5, 6, 8)
    print("", flush=True)
    print("| draw stream (cells per CU per clock) | " + " | ".join(f"{w} w/SIMD" for w in wps)
          + " |", flush=True)
    print("|---" * (len(wps) + 1) + "|", flush=True)
    row = [lgca_amd.probe_valu(lgca_amd.PROBE_STREAM, w, 1, 256)[0] for w in wps]
    print("| xorshift64* + column + deposit | " + " | ".join(f"{r:.2f}" for r in row) + " |",
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--bands", type=int, default=0, help="CPU bands (default: host threads, <= 16)")
    ap.add_argument("--cells-per-lane", type=int, default=64, choices=(64, 32),
                    help="step-kernel work unit: a whole u64 word or half of it")
    ap.add_argument("--passes", type=int, default=0, choices=(0, 1, 2),
                    help="kernels per step: draw pass + collide pass (2), one fused kernel (1) "
                         "or chosen by lattice size (0)")
    ap.add_argument("--draw-cells-per-lane", type=int, default=0, choices=(0, 64, 32),
                    help="work unit of the two-pass step's draw work (0: by lattice size)")
    ap.add_argument("--draw-ahead", type=int, default=-1, metavar="N",
                    help="two-pass step: draws of up to N steps per launch on a second stream "
                         "(0: draws share the collide launches; -1: by lattice size)")
    ap.add_argument("--temporal-depth", type=int, default=0, metavar="K",
                    help="collide K steps per launch on bands of rows held in registers "
                         "(lg_set_temporal_depth; 0: the engine's default)")
    ap.add_argument("--tile-rows", type=int, default=0, metavar="R",
                    help="valid rows per band of the temporally blocked collide (0: by size)")
    ap.add_argument("--resident", type=int, default=2, choices=(0, 1, 2),
                    help="resident bands: 0 off, 1 on, 2 by the engine's size rule")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--ensemble", type=int, default=256, metavar="N",
                    help="lattices in the batched-ensemble leg (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=3000)
    ap.add_argument("--serve", type=int, default=256, metavar="N",
                    help="requests in the steady-state serving leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=22.0,
                    help="arrivals per second of the steady-state leg (about 375 GCUPS offered, "
                         "under the batched server's capacity)")
    ap.add_argument("--serve-saturated", type=int, default=256, metavar="N",
                    help="requests in the saturated serving leg at --serve-saturated-rate "
                         "(0 skips it)")
    ap.add_argument("--serve-saturated-rate", type=float, default=50.0)
    ap.add_argument("--no-transport", action="store_true",
                    help="skip the transport-measure leg (the reference's --measure)")
    ap.add_argument("--no-frames", action="store_true",
                    help="skip the frame-output leg")
    ap.add_argument("--no-reference-suite", action="store_true",
                    help="skip the reference's own GPU benchmark cases")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the single-lattice autotune leg (lg_autotune on the headline workload)")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 512 x 256 leg (keeps profiles to the headline lattice)")
    ap.add_argument("--shear-steps", type=int, default=400, metavar="N",
                    help="steps of the 537M-cell shear-wave leg (0 skips it)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes (to scale rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the VALU issue-rate probe (lg_probe_valu) and print its table")
    args = ap.parse_args()

    if args.valu_probe:
        valu_probe_table()
        return

    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20)
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": 5}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this engine runs on exactly one GPU; --gpus {args.gpus} is not "
              "supported (lattices are not sharded across GPUs).", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    import numpy as np

    import workloads as W

    bands = args.bands or host_threads()
    cfg = W.config2(bands=bands)
    cells = cfg["w"] * cfg["h"]
    every = cfg["sample_every"]

    g = W.config_gpu(cfg)
    g.set_work_unit(args.cells_per_lane)
    g.set_passes(args.passes)
    g.set_draw_unit(args.draw_cells_per_lane)
    if args.draw_ahead >= 0:
        g.set_draw_ahead(args.draw_ahead)
    if args.temporal_depth:
        g.set_temporal_depth(args.temporal_depth, args.tile_rows)
    g.set_resident(args.resident)
    g.attach_field(cfg["block"], cfg["block"])
    if args.warmup:
        # through the same entry point as the timed call, as two calls: the
        # second call of a process still pays one-time host costs (~10 us on
        # a ~200 us call), which a single warm-up call would leave to the
        # timed one
        first = max(1, args.warmup - 1)
        g.advance_timed(first, every, 1.0)
        g.advance_timed(args.warmup - first, every, 1.0)

    # timed region: inputs resident in HBM; synchronous call on both sides
    t0 = time.perf_counter()
    dev_ms = g.advance_timed(args.steps, every, 1.0)
    wall = time.perf_counter() - t0

    gcups = cells * args.steps / wall / 1e9
    step_ms = dev_ms / args.steps
    traffic_alg = cells * BYTES_PER_CELL_UPDATE
    achieved = traffic_alg / (step_ms * 1e-3) / 1e9
    passes, cpl, dcpl, ahead = g.pipeline()  # what the timed steps ran
    bands_res = g.resident()
    unit = "h" if cpl == 32 else "w"
    nh, dnh = 3 - cpl // 32, 3 - dcpl // 32
    steps_per_launch = 1
    if bands_res:
        # one cooperative launch per chunk: a chunk ends at a sample (which
        # the launch itself does: the kernel's sampling instance) or at 65536
        # steps
        # the instance's inflow mode as launch_band picks it: 0 the table
        # wave draws the band's inflow cells, 1 each row's lanes, 2 the
        # serial per-word chain (inflow wider than 64 cells)
        cpr = 1
        while cpr < min(cfg["inlet"], 64):
            cpr *= 2
        rows = bands_res[0]
        inflow = 2 if cfg["inlet"] > 64 else (0 if -(-rows // (64 // cpr)) <= 3 else 1)
        kernel = (f"lg_band_kernel<{'true' if cfg['w'] > 4032 else 'false'},"
                  f"{'true' if every else 'false'},{inflow}>")
        steps_per_launch = min(args.steps, every or args.steps, 65536)
    elif passes == 1:
        kernel = f"lg_step_kernel_{unit}"
    elif ahead:
        kernel = f"lg_collide_kernel<{nh}>+lg_draw_kernel<{dnh}>"
    else:
        kernel = f"lg_step2_kernel<{nh},{dnh}>"
    launches = -(-args.steps // steps_per_launch)

    out = {
        "metric": METRIC,
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 6),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": cfg["name"], "w": cfg["w"], "h": cfg["h"], "cells": cells,
                   "rest": True, "bands": bands, "inlet_cols": cfg["inlet"],
                   "sample_every": every, "block": cfg["block"], "seed": cfg["seed"],
                   "cells_per_lane": cpl, "passes": passes,
                   "draw_cells_per_lane": dcpl if passes == 2 and not bands_res else None,
                   "draw_ahead": ahead if passes == 2 and not bands_res else None,
                   "resident_bands": ({"rows_per_band": bands_res[0], "bands": bands_res[1],
                                       "steps_per_launch": steps_per_launch}
                                      if bands_res else None)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(kernel, cfg["name"]),
                     "kernel": kernel,
                     "algorithmic_bytes_per_launch": int(traffic_alg * steps_per_launch),
                     "steps_per_launch": steps_per_launch,
                     "mean_launch_ms": round(dev_ms / launches, 6),
                     "mean_step_ms": round(step_ms, 6)},
    }
    if rank == 0 and not args.no_cpu:
        total = args.warmup + args.steps
        gpu_cells = g.store()
        out["cpu_baseline"] = cpu_baseline(cfg, total, gpu_cells, args.cpu_budget)
    g.close()
    try:
        if args.no_small:
            raise _Skip()
        out["small_512x256"], small_cfg, small_final = small_leg(
            bands, args.passes, args.cells_per_lane, args.draw_cells_per_lane, args.draw_ahead)
        if "cpu_baseline" in out:
            out["cpu_baseline"]["small_512x256"] = cpu_small(small_cfg, small_final)
    except _Skip:
        pass
    except Exception as e:
        out["small_512x256"] = {"error": repr(e)}
    if args.shear_steps:
        try:
            out["shear_537m"] = shear_leg(args.shear_steps, bands, args.passes, args.cells_per_lane,
                                          args.draw_cells_per_lane, args.draw_ahead)
        except Exception as e:
            out["shear_537m"] = {"error": repr(e)}
    if not args.no_transport:
        try:
            out["transport_measure"] = transport_leg()
        except Exception as e:
            out["transport_measure"] = {"error": repr(e)}
    if not args.no_frames:
        try:
            out["frames"], frame_field = frames_leg(bands, args.passes, args.cells_per_lane,
                                                    args.draw_cells_per_lane, args.draw_ahead)
            if "cpu_baseline" in out:
                out["cpu_baseline"]["frames"] = cpu_frames(frame_field)
        except Exception as e:
            out["frames"] = {"error": repr(e)}
    if not args.no_reference_suite:
        try:
            out["reference_suite"] = reference_suite_leg(bands)
        except Exception as e:
            out["reference_suite"] = {"error": repr(e)}
    if not args.no_autotune:
        try:
            out["autotune"] = autotune_leg(bands)
        except Exception as e:
            out["autotune"] = {"error": repr(e)}
    if args.ensemble:
        try:
            out["ensemble"] = ensemble_leg(args.ensemble, args.ensemble_steps, bands)
        except Exception as e:  # the headline stands on its own
            out["ensemble"] = {"error": repr(e)}
    if args.serve:
        try:
            out["serving"] = serving_leg(args.serve, args.serve_rate)
        except Exception as e:
            out["serving"] = {"error": repr(e)}
    if args.serve_saturated:
        try:
            out["serving_saturated"] = serving_leg(args.serve_saturated, args.serve_saturated_rate)
        except Exception as e:
            out["serving_saturated"] = {"error": repr(e)}
    try:  # the steady-state stream against the batched server's saturated throughput
        out["serving"]["offered_over_capacity"] = round(
            out["serving"]["offered_gcups"] / out["serving_saturated"]["batched"]["gcups"], 3)
    except (KeyError, TypeError, ZeroDivisionError):
        pass
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
