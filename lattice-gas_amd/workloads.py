"""The workloads the engine is measured and checked on, built on the device.

Sizes and statistics follow the benchmark configurations: the reference's
golden cases (tests/golden.rs:49-86), the 512x256 CPU-path run, and the
9.4M-cell default-scale run. Geometry and initialisation go through the
engine's own mirrors of Lattice::add_solid / init_equilibrium
(src/lattice.rs:197-220), so nothing here runs on the CPU but bookkeeping.
"""
from __future__ import annotations

import numpy as np

from lgca_amd import GpuLattice

SQRT3_2 = np.float32(0.8660254)  # src/hex.rs:25

# tests/golden.rs:49-74: (name, w, h, rest, inlet, plate, threads, steps)
GOLDEN_DENSITY, GOLDEN_SPEED, GOLDEN_SEED = 0.22, 0.3, 0x5EED
GOLDEN_CASES = [
    ("periodic/rest/t1", 128, 128, True, 0, False, 1, 200),
    ("periodic/no-rest/t1", 128, 128, False, 0, False, 1, 200),
    ("periodic/rest/t3", 128, 128, True, 0, False, 3, 200),
    ("channel/plate/inlet/t1", 128, 64, True, 8, True, 1, 200),
    ("channel/plate/inlet/t4", 128, 64, True, 8, True, 4, 200),
]


def f32(x):
    return np.float32(x)


def golden_plate(w: int, h: int):
    """tests/golden.rs:78-83 in f32: x in [w/5, w/5+3], y in yc +- 12."""
    yc = f32(f32(h) * SQRT3_2 / f32(2.0))
    x0 = f32(f32(w) / f32(5.0))
    return float(x0), float(f32(x0 + f32(3.0))), float(f32(yc - f32(12.0))), float(f32(yc + f32(12.0)))


def golden_gpu(case) -> GpuLattice:
    """tests/golden.rs:76-86 on the engine."""
    name, w, h, rest, inlet, plate, threads, steps = case
    g = GpuLattice(w, h, rest, GOLDEN_SEED, threads)
    g.set_inlet(inlet, GOLDEN_DENSITY, (GOLDEN_SPEED, 0.0))
    if plate:
        g.add_solid_box(*golden_plate(w, h))
    g.init_equilibrium(GOLDEN_DENSITY, GOLDEN_SPEED, 0.0)
    return g


def plate_box(w: int, h: int, x0: float, rows: int):
    """The reference's plate (src/main.rs:582-588): x in [x0, x0 + 4], centred
    at yc = h * sqrt(3)/2 / 2, `rows` rows tall (size = rows * sqrt(3)/2)."""
    yc = f32(f32(h) * SQRT3_2 / f32(2.0))
    size = f32(f32(rows) * SQRT3_2)
    half = f32(size / f32(2.0))
    return (float(f32(x0)), float(f32(f32(x0) + f32(4.0))), float(f32(yc - half)), float(f32(yc + half)))


def config1(bands: int = 1) -> dict:
    """512x256 hex lattice, rest particles, links and rest Bernoulli(0.25)
    from seed 42, plate 64 rows tall at x = 128, 8 inflow columns at drift
    0.2, 2000 steps, 16x16 block sums."""
    return dict(name="fhp-512x256-plate", w=512, h=256, rest=True, density=0.25, init_u=(0.0, 0.0),
                inflow=(0.2, 0.0), inlet=8, seed=42, bands=bands,
                plate=plate_box(512, 256, 128.0, 64), steps=2000, block=16)


def config2(bands: int = 1) -> dict:
    """~9.4M cells (4096x2304), same statistics, plate H/8 tall at x = W/4,
    5000 steps, 32x32 coarse-grain every 100 steps."""
    return dict(name="fhp-4096x2304-plate", w=4096, h=2304, rest=True, density=0.25,
                init_u=(0.0, 0.0), inflow=(0.2, 0.0), inlet=8, seed=42, bands=bands,
                plate=plate_box(4096, 2304, 1024.0, 2304 // 8), steps=5000, block=32,
                sample_every=100)


def render_bench(bands: int = 1) -> dict:
    """The field the reference times its frame writers on (benches/speed.rs:
    330-357, 633-657): 2048 x 1280, rest particles, density 0.22 at drift 0.4
    (also the 8 inflow columns'), seed 0x5EED, a plate 4 wide at x = w / 5 and
    h * 173 / 1280 tall about yc, 50 steps, blocks of 20 sampled at alpha 1."""
    w, h = 2048, 1280
    yc = f32(f32(h) * SQRT3_2 / f32(2.0))
    size = f32(f32(h) * f32(173.0) / f32(1280.0))
    x0 = f32(f32(w) / f32(5.0))
    half = f32(size / f32(2.0))
    plate = (float(x0), float(f32(x0 + f32(4.0))), float(f32(yc - half)), float(f32(yc + half)))
    return dict(name="fhp-2048x1280-render", w=w, h=h, rest=True, density=0.22,
                init_u=(0.4, 0.0), inflow=(0.4, 0.0), inlet=8, seed=0x5EED, bands=bands,
                plate=plate, steps=50, block=20)


def shear_profile(w: int, amp: float = 0.05):
    """u_y(x) = amp * sin(2 pi x / W) in f32, as src/transport.rs:234-242
    evaluates its column profile (k = 2 pi / W in f32, then k * x)."""
    k = np.float32(np.float32(2.0) * np.float32(np.pi) / np.float32(w))
    x = np.arange(w, dtype=np.float32)
    return (np.float32(amp) * np.sin((k * x).astype(np.float32))).astype(np.float32)


def config3(bands: int = 1, w: int = 32768, h: int = 16384) -> dict:
    """Transport-measure workload: periodic w x h lattice (default 32768 x
    16384, ~537M cells), no obstacle, no inflow, density 0.3 per link with
    rest particles, transverse shear wave u_y = 0.05 sin(2 pi x / W) imposed by
    biased link occupancy from seed 7 (init stream of src/lattice.rs:197-208,
    one equilibrium per column), 4000 steps."""
    return dict(name=f"fhp-{w}x{h}-shear", w=w, h=h, rest=True, density=0.3, init_u=(0.0, 0.0),
                inflow=(0.0, 0.0), inlet=0, seed=7, bands=bands, plate=None, steps=4000,
                profile="shear")


def config4(n: int = 256, bands: int = 16, seed: int = 4, steps: int = 3000) -> list:
    """Batched ensemble: n independent lattices, widths and heights uniform on
    256..2048 in steps of 64, a seed per lattice, rest particles at density
    0.25, a plate 10-25% of H tall at x uniform in [W/8, W/3], 8 inflow
    columns at a drift uniform in [0.1, 0.3]; `steps` steps each. One config
    per lattice (config_gpu builds it)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        w = int(rng.integers(4, 33)) * 64
        h = int(rng.integers(4, 33)) * 64
        lseed = int(rng.integers(1, 2**62))
        rows = max(2, int(round(rng.uniform(0.10, 0.25) * h)))
        x = float(rng.uniform(w / 8.0, w / 3.0))
        drift = float(rng.uniform(0.1, 0.3))
        out.append(dict(name=f"fhp-ensemble-{i}-{w}x{h}", w=w, h=h, rest=True, density=0.25,
                        init_u=(0.0, 0.0), inflow=(drift, 0.0), inlet=8, seed=lseed, bands=bands,
                        plate=plate_box(w, h, x, rows), steps=steps))
    return out


def config_gpu(cfg: dict) -> GpuLattice:
    g = GpuLattice(cfg["w"], cfg["h"], cfg["rest"], cfg["seed"], cfg["bands"])
    g.set_inlet(cfg["inlet"], cfg["density"], cfg["inflow"])
    if cfg.get("plate"):
        g.add_solid_box(*cfg["plate"])
    if cfg.get("profile") == "shear":
        uy = shear_profile(cfg["w"])
        g.init_profile(0, cfg["density"], np.zeros_like(uy), uy)
    else:
        g.init_equilibrium(cfg["density"], *cfg["init_u"])
    return g


def exact_invariants(cells: np.ndarray):
    """Exact (mass, px2, py2) over every cell (tests/golden.rs:90-105)."""
    c = np.asarray(cells, dtype=np.uint8)
    cnt = [int(np.count_nonzero(c & (1 << d))) for d in range(7)]
    cx2 = [2, 1, -1, -2, -1, 1]
    cy2 = [0, 1, 1, 0, -1, -1]
    mass = sum(cnt)
    return mass, sum(a * b for a, b in zip(cx2, cnt[:6])), sum(a * b for a, b in zip(cy2, cnt[:6]))
