"""ctypes binding of the CPU oracle (oracle/lgca_oracle.hpp).

TEST INFRASTRUCTURE ONLY. The oracle is what the HIP engine's results are
compared with; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and never as the thing measured or shipped.

The wrapper mirrors the reference's own types: ``Lattice`` restates
src/lattice.rs:8-364 and ``Field`` restates src/render.rs:21-120.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liblgca_oracle.so")
_lib = None

SQRT3_2 = np.float32(0.8660254)  # src/hex.rs:25


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc; no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, sz, u64, f32, i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_float, C.c_int
        P = C.POINTER
        sig = {
            "orc_new": (vp, [sz, sz, f32, f32, f32, sz, i32, sz, u64]),
            "orc_free": (None, [vp]),
            "orc_set_parallel": (None, [vp, i32]),
            "orc_add_solid_box": (None, [vp, f32, f32, f32, f32]),
            "orc_add_solid_disc": (None, [vp, f32, f32, f32]),
            "orc_init_equilibrium": (None, [vp]),
            "orc_set_inflow": (None, [vp, f32, f32]),
            "orc_init_profile": (None, [vp, i32, P(f32), P(f32)]),
            "orc_restart": (i32, [vp, P(C.c_uint8), sz, u64]),
            "orc_advance": (None, [vp, u64]),
            "orc_steps": (u64, [vp]),
            "orc_n_bands": (sz, [vp]),
            "orc_get_cells": (i32, [vp, P(C.c_uint8), sz]),
            "orc_set_cells": (i32, [vp, P(C.c_uint8), sz]),
            "orc_total_particles": (u64, [vp]),
            "orc_mean_velocity": (None, [vp, P(f32), P(f32)]),
            "orc_active_states": (sz, [vp]),
            "orc_field_new": (vp, [sz, sz, sz, sz]),
            "orc_field_free": (None, [vp]),
            "orc_field_dims": (None, [vp, P(sz), P(sz)]),
            "orc_field_sample": (i32, [vp, vp, f32]),
            "orc_field_read": (None, [vp, P(f32), P(f32), P(f32), P(f32)]),
            "orc_field_checksum": (u64, [vp]),
            "orc_block_sums": (None, [vp, vp] + [P(C.c_int64)] * 4),
            "orc_fnv1a": (u64, [P(C.c_uint8), sz]),
            "orc_thresholds": (None, [f32, f32, f32, i32, P(C.c_uint32)]),
            "orc_eq_samples": (None, [f32, f32, f32, i32, u64, sz, P(C.c_uint8)]),
            "orc_rng_new": (u64, [u64]),
            "orc_rng_u64": (None, [u64, sz, P(u64)]),
            "orc_table": (sz, [i32, P(C.c_uint8), sz]),
            "orc_classes": (sz, [i32, P(C.c_uint8), P(C.c_uint32)]),
            "orc_moment_tables": (None, [P(u64), P(u64)]),
            "orc_moments_sum": (None, [i32, P(C.c_uint8), sz, P(C.c_int64), P(f32)]),
            "orc_viscosity": (f32, [f32, i32, sz, u64, sz]),
            "orc_advection_factor": (f32, [f32, i32, sz, u64, sz]),
            "orc_row_ux": (None, [P(C.c_uint8), sz, sz, P(f32)]),
            "orc_column_uy": (None, [P(C.c_uint8), sz, sz, P(f32)]),
            "orc_vorticity": (None, [P(f32), P(f32), sz, sz, sz, sz, P(f32)]),
            "orc_vorticity_scale": (f32, [P(f32), P(f32), sz]),
            "orc_field_mean_velocity": (None, [P(f32), P(f32), P(f32), sz, P(f32), P(f32)]),
            "orc_deflate": (sz, [P(C.c_uint8), sz, P(C.c_uint8), sz]),
            "orc_crc32": (C.c_uint32, [P(C.c_uint8), sz]),
            "orc_adler32": (C.c_uint32, [P(C.c_uint8), sz]),
            "orc_png_rgb": (sz, [sz, sz, P(C.c_uint8), P(C.c_uint8), sz]),
            "orc_vorticity_png": (sz, [P(f32), P(f32), P(f32), sz, sz, sz, sz, sz, f32,
                                       P(C.c_uint8), sz]),
            "orc_arrows_svg": (sz, [P(f32), P(f32), P(f32), sz, sz, sz, sz, f32, f32, f32,
                                    C.c_char_p, sz]),
            "orc_fmt1": (sz, [f32, C.c_char_p, sz]),
            "orc_fmt_short": (sz, [f32, C.c_char_p, sz]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ---- free functions ---------------------------------------------------------

def fnv1a(b) -> int:
    a = np.ascontiguousarray(np.frombuffer(bytes(b), dtype=np.uint8) if not isinstance(b, np.ndarray) else b.view(np.uint8).ravel())
    return int(lib().orc_fnv1a(_ptr(a, C.c_uint8), a.size))


def thresholds(density: float, ux: float, uy: float, rest: bool) -> np.ndarray:
    out = np.zeros(7, dtype=np.uint32)
    lib().orc_thresholds(density, ux, uy, int(rest), _ptr(out, C.c_uint32))
    return out


def eq_samples(density, ux, uy, rest, seed, n) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint8)
    lib().orc_eq_samples(density, ux, uy, int(rest), seed, n, _ptr(out, C.c_uint8))
    return out


def rng_state(seed: int) -> int:
    return int(lib().orc_rng_new(seed))


def rng_u64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_rng_u64(seed, n, _ptr(out, C.c_uint64))
    return out


# ---- frame output (oracle/lgca_render.hpp) -----------------------------------

def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).ravel()


def _bytes_call(fn, *args) -> bytes:
    n = fn(*args, None, 0)
    out = C.create_string_buffer(max(n, 1))
    fn(*args, C.cast(out, fn.argtypes[-2]), n)
    return out.raw[:n]


def vorticity(ux, uy, bw, bh, bx, by) -> np.ndarray:
    """Field::vorticity (src/render.rs:122-139)."""
    ux, uy = _f32(ux), _f32(uy)
    out = np.zeros(bw * bh, dtype=np.float32)
    lib().orc_vorticity(_ptr(ux, C.c_float), _ptr(uy, C.c_float), bw, bh, bx, by, _ptr(out, C.c_float))
    return out


def vorticity_scale(vort, solid) -> float:
    """write_frame's colour scale (src/main.rs:219-229)."""
    v, s = _f32(vort), _f32(solid)
    return float(lib().orc_vorticity_scale(_ptr(v, C.c_float), _ptr(s, C.c_float), v.size))


def field_mean_velocity(ux, uy, solid):
    """Field::mean_velocity (src/render.rs:142-157)."""
    ux, uy, s = _f32(ux), _f32(uy), _f32(solid)
    mx, my = C.c_float(), C.c_float()
    lib().orc_field_mean_velocity(_ptr(ux, C.c_float), _ptr(uy, C.c_float), _ptr(s, C.c_float),
                                  ux.size, C.byref(mx), C.byref(my))
    return mx.value, my.value


def _u8buf(data) -> np.ndarray:
    return np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))


def deflate(data) -> bytes:
    """png::deflate (src/png.rs:103-162)."""
    a = _u8buf(data)
    return _bytes_call(lib().orc_deflate, _ptr(a, C.c_uint8), a.size)


def crc32(data) -> int:
    a = _u8buf(data)
    return int(lib().orc_crc32(_ptr(a, C.c_uint8), a.size))


def adler32(data) -> int:
    a = _u8buf(data)
    return int(lib().orc_adler32(_ptr(a, C.c_uint8), a.size))


def png_rgb(rgb: np.ndarray) -> bytes:
    """png::write_rgb (src/png.rs:174-197) of an (h, w, 3) u8 image: the file's bytes."""
    img = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w = img.shape[:2]
    return _bytes_call(lib().orc_png_rgb, w, h, _ptr(img, C.c_uint8))


def vorticity_png(ux, uy, solid, bw, bh, bx, by, scale, wmax) -> bytes:
    """render::write_vorticity (src/render.rs:176-204): the file's bytes."""
    ux, uy, s = _f32(ux), _f32(uy), _f32(solid)
    return _bytes_call(lib().orc_vorticity_png, _ptr(ux, C.c_float), _ptr(uy, C.c_float),
                       _ptr(s, C.c_float), bw, bh, bx, by, scale, wmax)


def arrows_svg(ux, uy, solid, bw, bh, bx, by, scale, frame) -> bytes:
    """render::write_arrows (src/render.rs:209-267): the file's bytes."""
    ux, uy, s = _f32(ux), _f32(uy), _f32(solid)
    return _bytes_call(lib().orc_arrows_svg, _ptr(ux, C.c_float), _ptr(uy, C.c_float),
                       _ptr(s, C.c_float), bw, bh, bx, by, scale, frame[0], frame[1])


def fmt1(x) -> str:
    """Rust's `{:.1}` of an f32."""
    return _bytes_call(lib().orc_fmt1, float(np.float32(x))).rstrip(b"\0").decode()


def fmt_short(x) -> str:
    """Rust's `{}` of an f32."""
    return _bytes_call(lib().orc_fmt_short, float(np.float32(x))).rstrip(b"\0").decode()


def collision_table(rest: bool):
    n_states = 128 if rest else 64
    stride = lib().orc_table(int(rest), None, 0)
    out = np.zeros(n_states * stride, dtype=np.uint8)
    lib().orc_table(int(rest), _ptr(out, C.c_uint8), out.size)
    return stride, out.reshape(n_states, stride)


def classes(rest: bool):
    members = np.zeros(128, dtype=np.uint8)
    sizes = np.zeros(128, dtype=np.uint32)
    n = lib().orc_classes(int(rest), _ptr(members, C.c_uint8), _ptr(sizes, C.c_uint32))
    out, off = [], 0
    for c in range(n):
        s = int(sizes[c])
        out.append([int(v) for v in members[off:off + s]])
        off += s
    return out


def moment_tables():
    f = np.zeros(256, dtype=np.uint64)
    w = np.zeros(256, dtype=np.uint64)
    lib().orc_moment_tables(_ptr(f, C.c_uint64), _ptr(w, C.c_uint64))
    return f, w


def moments_sum(cells: np.ndarray, walls: bool):
    a = np.ascontiguousarray(cells, dtype=np.uint8).ravel()
    out = np.zeros(4, dtype=np.int64)
    v = np.zeros(2, dtype=np.float32)
    lib().orc_moments_sum(int(walls), _ptr(a, C.c_uint8), a.size, _ptr(out, C.c_int64), _ptr(v, C.c_float))
    return tuple(int(x) for x in out), (float(v[0]), float(v[1]))


def viscosity(density, rest, threads, seed, size=256) -> float:
    """transport::viscosity (src/transport.rs:183-221)."""
    return float(lib().orc_viscosity(density, int(rest), threads, seed, size))


def advection_factor(density, rest, threads, seed, size=256) -> float:
    """transport::advection_factor (src/transport.rs:225-289)."""
    return float(lib().orc_advection_factor(density, int(rest), threads, seed, size))


def row_ux(cells, w, h):
    a = np.ascontiguousarray(cells, dtype=np.uint8)
    out = np.zeros(h, dtype=np.float32)
    lib().orc_row_ux(_ptr(a, C.c_uint8), w, h, _ptr(out, C.c_float))
    return out


def column_uy(cells, w, h):
    a = np.ascontiguousarray(cells, dtype=np.uint8)
    out = np.zeros(w, dtype=np.float32)
    lib().orc_column_uy(_ptr(a, C.c_uint8), w, h, _ptr(out, C.c_float))
    return out


# ---- Lattice / Field -------------------------------------------------------

class Lattice:
    """src/lattice.rs:8-364, byte per cell."""

    def __init__(self, w, h, density, inflow, inlet_cols, rest, threads, seed):
        self.w, self.h = int(w), int(h)
        self.rest = bool(rest)
        self.density = float(density)
        self.inflow = (float(inflow[0]), float(inflow[1]))
        self.inlet_cols = int(inlet_cols)
        self.threads = max(int(threads), 1)
        self.seed = int(seed)
        self._p = lib().orc_new(self.w, self.h, self.density, self.inflow[0], self.inflow[1],
                                self.inlet_cols, int(self.rest), self.threads, self.seed & (2**64 - 1))
        if not self._p:
            raise ValueError("row count must be even and w, h > 4 (src/lattice.rs:137-138)")

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            lib().orc_free(p)
            self._p = None

    def set_parallel(self, on: bool):
        lib().orc_set_parallel(self._p, int(on))

    def add_solid_box(self, x0, x1, y0, y1):
        lib().orc_add_solid_box(self._p, x0, x1, y0, y1)

    def add_solid_disc(self, cx, cy, r):
        lib().orc_add_solid_disc(self._p, cx, cy, r)

    def init_equilibrium(self):
        lib().orc_init_equilibrium(self._p)

    def init_profile(self, axis, ux, uy):
        a = np.ascontiguousarray(ux, dtype=np.float32)
        b = np.ascontiguousarray(uy, dtype=np.float32)
        lib().orc_init_profile(self._p, int(axis), _ptr(a, C.c_float), _ptr(b, C.c_float))

    def set_inflow(self, ux, uy):
        lib().orc_set_inflow(self._p, ux, uy)
        self.inflow = (float(ux), float(uy))

    def restart(self, cells: np.ndarray, seed: int):
        a = np.ascontiguousarray(cells, dtype=np.uint8).ravel()
        if lib().orc_restart(self._p, _ptr(a, C.c_uint8), a.size, seed & (2**64 - 1)) != 0:
            raise ValueError("wrong number of cells")
        self.seed = int(seed)

    def advance(self, n: int):
        lib().orc_advance(self._p, int(n))

    @property
    def steps(self) -> int:
        return int(lib().orc_steps(self._p))

    @property
    def n_bands(self) -> int:
        return int(lib().orc_n_bands(self._p))

    @property
    def cells(self) -> np.ndarray:
        out = np.zeros(self.w * self.h, dtype=np.uint8)
        lib().orc_get_cells(self._p, _ptr(out, C.c_uint8), out.size)
        return out

    @cells.setter
    def cells(self, v: np.ndarray):
        a = np.ascontiguousarray(v, dtype=np.uint8).ravel()
        if lib().orc_set_cells(self._p, _ptr(a, C.c_uint8), a.size) != 0:
            raise ValueError("wrong number of cells")

    def total_particles(self) -> int:
        return int(lib().orc_total_particles(self._p))

    def mean_velocity(self):
        ux, uy = C.c_float(), C.c_float()
        lib().orc_mean_velocity(self._p, C.byref(ux), C.byref(uy))
        return ux.value, uy.value

    def active_states(self) -> int:
        return int(lib().orc_active_states(self._p))


class Field:
    """src/render.rs:21-120."""

    def __init__(self, w, h, bx, by):
        self._p = lib().orc_field_new(w, h, bx, by)
        if not self._p:
            raise ValueError("block size must be positive")
        bw, bh = C.c_size_t(), C.c_size_t()
        lib().orc_field_dims(self._p, C.byref(bw), C.byref(bh))
        self.bw, self.bh, self.bx, self.by = bw.value, bh.value, bx, by

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            lib().orc_field_free(p)
            self._p = None

    def sample(self, lat: Lattice, alpha: float):
        if lib().orc_field_sample(self._p, lat._p, alpha) != 0:
            raise ValueError("block larger than MAX_BLOCK (src/render.rs:62-70)")

    def read(self):
        n = self.bw * self.bh
        arrs = [np.zeros(n, dtype=np.float32) for _ in range(4)]
        lib().orc_field_read(self._p, *[_ptr(a, C.c_float) for a in arrs])
        return tuple(arrs)

    def checksum(self) -> int:
        return int(lib().orc_field_checksum(self._p))

    def block_sums(self, lat: Lattice):
        n = self.bw * self.bh
        arrs = [np.zeros(n, dtype=np.int64) for _ in range(4)]
        lib().orc_block_sums(self._p, lat._p, *[_ptr(a, C.c_int64) for a in arrs])
        return tuple(arrs)


# ---- tests/golden.rs:49-132 ------------------------------------------------

GOLDEN_DENSITY = 0.22
GOLDEN_SPEED = 0.3
GOLDEN_SEED = 0x5EED
GOLDEN_BLOCK = 16

# tests/golden.rs:68-74: (name, w, h, rest, inlet, plate, threads, steps)
GOLDEN_CASES = [
    ("periodic/rest/t1", 128, 128, True, 0, False, 1, 200),
    ("periodic/no-rest/t1", 128, 128, False, 0, False, 1, 200),
    ("periodic/rest/t3", 128, 128, True, 0, False, 3, 200),
    ("channel/plate/inlet/t1", 128, 64, True, 8, True, 1, 200),
    ("channel/plate/inlet/t4", 128, 64, True, 8, True, 4, 200),
]


def golden_plate(w: int, h: int):
    """tests/golden.rs:78-83: x in [w/5, w/5+3], y in [yc-12, yc+12], f32."""
    yc = np.float32(np.float32(h) * SQRT3_2 / np.float32(2.0))
    x0 = np.float32(np.float32(w) / np.float32(5.0))
    x1 = np.float32(x0 + np.float32(3.0))
    return float(x0), float(x1), float(np.float32(yc - np.float32(12.0))), float(np.float32(yc + np.float32(12.0)))


def golden_build(case, seed=GOLDEN_SEED) -> Lattice:
    """tests/golden.rs:76-86."""
    name, w, h, rest, inlet, plate, threads, steps = case
    lat = Lattice(w, h, GOLDEN_DENSITY, (GOLDEN_SPEED, 0.0), inlet, rest, threads, seed)
    if plate:
        lat.add_solid_box(*golden_plate(w, h))
    lat.init_equilibrium()
    return lat


def exact_invariants(cells: np.ndarray):
    """tests/golden.rs:90-105: exact (mass, px2, py2) over every cell."""
    cx2 = np.array([2, 1, -1, -2, -1, 1], dtype=np.int64)
    cy2 = np.array([0, 1, 1, 0, -1, -1], dtype=np.int64)
    c = cells.astype(np.uint8)
    bits = np.unpackbits(c[:, None], axis=1, bitorder="little").astype(np.int64)
    mass = int(bits[:, :7].sum())
    px = int((bits[:, :6] * cx2).sum())
    py = int((bits[:, :6] * cy2).sum())
    return mass, px, py
