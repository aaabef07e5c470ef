"""Python host side of the MI355X FHP engine, over the C ABI of include/lgca.h.

``GpuLattice`` mirrors the reference's ``GpuLattice`` (src/gpu.rs:453-755)
method for method, and ``Sim`` mirrors ``Sim`` (src/sim.rs:19-175) for its GPU
backend. Every call goes to liblgca_hip.so; there is no CPU fallback in this
module: if the library or a HIP device is missing, construction raises.

Cells are one byte per site in the reference's layout (src/hex.rs:7-23):
bits 0-5 moving directions, bit 6 rest, bit 7 solid, row-major ``y * w + x``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblgca_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "lgca.h")

_lib = None


class LgcaError(RuntimeError):
    pass


def load_library(path: str | None = None):
    """Load liblgca_hip.so (or the library named by $LGCA_LIB) and declare its
    C signatures (no device calls)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("LGCA_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise LgcaError(f"{path} is missing: build it with lattice-gas_amd/build.py")
    L = C.CDLL(path)
    vp, sz, u64, f32, i32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_float, C.c_int
    P = C.POINTER
    sig = {
        "lg_create": (vp, [sz, sz, i32, u64, sz]),
        "lg_destroy": (None, [vp]),
        "lg_last_error": (C.c_char_p, []),
        "lg_trim_memory": (i32, [P(sz)]),
        "lg_device_name": (i32, [vp, C.c_char_p, sz]),
        "lg_set_inlet": (i32, [vp, sz, P(C.c_uint32)]),
        "lg_set_inlet_flow": (i32, [vp, sz, f32, f32, f32]),
        "lg_attach_field": (i32, [vp, sz, sz]),
        "lg_load": (i32, [vp, P(C.c_uint8), sz]),
        "lg_restart": (i32, [vp, P(C.c_uint8), sz, u64]),
        "lg_store": (i32, [vp, P(C.c_uint8), sz]),
        "lg_advance": (i32, [vp, u64]),
        "lg_advance_sampling": (i32, [vp, u64, u64, f32]),
        "lg_advance_async": (i32, [vp, u64]),
        "lg_advance_sampling_async": (i32, [vp, u64, u64, f32]),
        "lg_poll": (i32, [vp]),
        "lg_wait": (i32, [vp]),
        "lg_sample_field": (i32, [vp, f32]),
        "lg_read_field": (i32, [vp, P(f32), P(f32), P(f32), P(f32), sz]),
        "lg_read_block_sums": (i32, [vp] + [P(C.c_int64)] * 4 + [sz]),
        "lg_total_particles": (i32, [vp, P(u64)]),
        "lg_dims": (i32, [vp, P(sz), P(sz)]),
        "lg_steps": (u64, [vp]),
        "lg_set_steps": (i32, [vp, u64]),
        "lg_set_work_unit": (i32, [vp, C.c_uint32]),
        "lg_set_passes": (i32, [vp, C.c_uint32]),
        "lg_set_draw_unit": (i32, [vp, C.c_uint32]),
        "lg_set_draw_ahead": (i32, [vp, C.c_uint32]),
        "lg_set_temporal_depth": (i32, [vp, C.c_uint32, C.c_uint32]),
        "lg_set_resident": (i32, [vp, C.c_uint32]),
        "lg_get_resident": (i32, [vp, P(C.c_uint32), P(C.c_uint32)]),
        "lg_get_temporal_depth": (i32, [vp, P(C.c_uint32), P(C.c_uint32)]),
        "lg_autotune": (i32, [vp, u64, P(u64)]),
        "lg_get_pipeline": (i32, [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "lg_batch_set_units": (i32, [vp, C.c_uint32]),
        "lg_batch_get_units": (i32, [vp, P(C.c_uint32)]),
        "lg_batch_set_waves": (i32, [vp, C.c_uint32]),
        "lg_batch_get_waves": (i32, [vp, P(C.c_uint32)]),
        "lg_batch_set_passes": (i32, [vp, C.c_uint32]),
        "lg_batch_set_draw_unit": (i32, [vp, C.c_uint32]),
        "lg_batch_get_pipeline": (i32, [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "lg_advance_timed": (i32, [vp, u64, P(f32)]),
        "lg_advance_sampling_timed": (i32, [vp, u64, u64, f32, P(f32)]),
        "lg_debug_collide_table": (i32, [i32, P(C.c_uint8), sz]),
        "lg_probe_op_name": (C.c_char_p, [i32]),
        "lg_probe_valu": (i32, [i32, C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_double),
                                P(C.c_double)]),
        "lg_add_solid_box": (i32, [vp, f32, f32, f32, f32]),
        "lg_add_solid_disc": (i32, [vp, f32, f32, f32]),
        "lg_init_equilibrium": (i32, [vp, f32, f32, f32]),
        "lg_init_profile": (i32, [vp, i32, f32, P(f32), P(f32), sz]),
        "lg_batch_create": (vp, [P(vp), sz]),
        "lg_batch_destroy": (None, [vp]),
        "lg_batch_advance": (i32, [vp, u64]),
        "lg_batch_set_shape": (i32, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
        "lg_batch_get_shape": (i32, [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
        "lg_batch_autotune": (i32, [vp, u64, P(u64)]),
        "lg_batch_sample": (i32, [vp, f32]),
        "lg_batch_advance_limited": (i32, [vp, u64, P(u64)]),
        "lg_batch_advance_async": (i32, [vp, u64, P(u64)]),
        "lg_batch_poll": (i32, [vp]),
        "lg_batch_wait": (i32, [vp]),
        "lg_batch_advance_sampling_async": (i32, [vp, u64, P(u64), f32]),
        "lg_batch_read_snapshot": (i32, [vp, P(f32), sz]),
        "lg_batch_snapshot": (i32, [vp, f32, P(f32), sz]),
        "lg_batch_snapshot_size": (sz, [vp]),
        "lg_batch_snapshot_noalloc": (i32, [vp, f32, P(f32), sz]),
        "lg_batch_snapshot_noalloc_max": (sz, []),
        "lg_sample_field_async": (i32, [vp, f32]),
        "lg_measure_viscosity": (i32, [f32, i32, sz, u64, sz, P(f32)]),
        "lg_measure_advection": (i32, [f32, i32, sz, u64, sz, P(f32)]),
        "lg_measure_transport": (i32, [f32, i32, sz, u64, sz, P(f32), P(f32)]),
        "lg_field_dims": (i32, [vp, P(sz), P(sz), P(sz), P(sz)]),
        "lg_field_vorticity": (i32, [P(f32), P(f32), sz, sz, sz, sz, P(f32)]),
        "lg_field_mean_velocity": (i32, [P(f32), P(f32), P(f32), sz, P(f32), P(f32)]),
        "lg_vorticity_scale": (i32, [P(f32), P(f32), sz, P(f32)]),
        "lg_png_deflate": (i32, [P(C.c_uint8), sz, P(C.c_uint8), sz, P(sz)]),
        "lg_write_png_rgb": (i32, [C.c_char_p, sz, sz, P(C.c_uint8)]),
        "lg_write_vorticity": (i32, [C.c_char_p, P(f32), P(f32), P(f32), sz, sz, sz, sz, sz, f32]),
        "lg_write_arrows": (i32, [C.c_char_p, P(f32), P(f32), P(f32), sz, sz, sz, sz, f32, f32, f32]),
        "lg_frames_create": (vp, [sz]),
        "lg_frames_write": (i32, [vp, P(f32), P(f32), P(f32), sz, sz, sz, sz, C.c_char_p,
                                  C.c_char_p, sz, f32, f32, f32]),
        "lg_frames_write_lattice": (i32, [vp, vp, C.c_char_p, C.c_char_p, sz, f32]),
        "lg_frames_finish": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(rc: int):
    if rc < 0:
        raise LgcaError(load_library().lg_last_error().decode())
    return rc


def _u8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


class GpuLattice:
    """src/gpu.rs:453-755, exact with the CPU path for the given band count."""

    def __init__(self, w: int, h: int, rest_particles: bool, seed: int, bands: int = 1):
        L = load_library()
        self._p = L.lg_create(int(w), int(h), int(bool(rest_particles)), int(seed) & (2**64 - 1),
                              max(int(bands), 1))
        if not self._p:
            raise LgcaError(L.lg_last_error().decode())
        self.w, self.h = int(w), int(h)
        self.rest = bool(rest_particles)
        self.bands = max(int(bands), 1)
        self._field = None

    def close(self):
        p = getattr(self, "_p", None)
        if p:
            load_library().lg_destroy(p)
            self._p = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # src/gpu.rs:517-519
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        _check(load_library().lg_device_name(self._p, buf, 256))
        return buf.value.decode()

    # src/gpu.rs:525-528
    def set_inlet(self, cols: int, density: float, inflow, rest: bool | None = None):
        if rest is not None and bool(rest) != self.rest:
            raise LgcaError("rest-particle setting differs from the lattice's")
        _check(load_library().lg_set_inlet_flow(self._p, int(cols), float(density),
                                                float(inflow[0]), float(inflow[1])))

    def set_inlet_thresholds(self, cols: int, thresh7):
        t = np.ascontiguousarray(thresh7, dtype=np.uint32)
        _check(load_library().lg_set_inlet(self._p, int(cols), t.ctypes.data_as(C.POINTER(C.c_uint32))))

    # src/gpu.rs:532-545
    def attach_field(self, bx: int, by: int):
        _check(load_library().lg_attach_field(self._p, int(bx), int(by)))
        self._field = (int(bx), int(by), self.w // int(bx), self.h // int(by))

    # src/lattice.rs:210-220 (box / disc obstacles), on the device
    def add_solid_box(self, x0, x1, y0, y1):
        _check(load_library().lg_add_solid_box(self._p, float(x0), float(x1), float(y0), float(y1)))

    def add_solid_disc(self, cx, cy, r):
        _check(load_library().lg_add_solid_disc(self._p, float(cx), float(cy), float(r)))

    # src/lattice.rs:197-208, on the device
    def init_equilibrium(self, density, ux, uy):
        _check(load_library().lg_init_equilibrium(self._p, float(density), float(ux), float(uy)))

    def init_profile(self, axis: int, density: float, ux, uy):
        """init_equilibrium with a per-column (axis 0) or per-row (axis 1)
        velocity profile."""
        a = np.ascontiguousarray(ux, dtype=np.float32)
        b = np.ascontiguousarray(uy, dtype=np.float32)
        _check(load_library().lg_init_profile(self._p, int(axis), float(density),
                                              a.ctypes.data_as(C.POINTER(C.c_float)),
                                              b.ctypes.data_as(C.POINTER(C.c_float)), a.size))

    # src/gpu.rs:572-610
    def load(self, cells: np.ndarray):
        a = np.ascontiguousarray(cells, dtype=np.uint8).ravel()
        _check(load_library().lg_load(self._p, _u8(a), a.size))

    # src/gpu.rs:616-620
    def restart(self, cells: np.ndarray, seed: int):
        a = np.ascontiguousarray(cells, dtype=np.uint8).ravel()
        _check(load_library().lg_restart(self._p, _u8(a), a.size, int(seed) & (2**64 - 1)))

    # src/gpu.rs:630-652
    def store(self, out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.w * self.h, dtype=np.uint8)
        _check(load_library().lg_store(self._p, _u8(out), out.size))
        return out

    # src/gpu.rs:657-659
    def advance(self, n: int):
        _check(load_library().lg_advance(self._p, int(n)))

    # src/gpu.rs:668-706
    def advance_sampling(self, n: int, sample_every: int, alpha: float):
        _check(load_library().lg_advance_sampling(self._p, int(n), int(sample_every), float(alpha)))

    def advance_async(self, n: int, sample_every: int = 0, alpha: float = 0.0):
        """advance / advance_sampling without the wait: returns once the steps
        are queued; later calls on the lattice queue behind them."""
        _check(load_library().lg_advance_sampling_async(self._p, int(n), int(sample_every),
                                                         float(alpha)))

    def poll(self) -> bool:
        """True once the lattice's queued work has ended."""
        return bool(_check(load_library().lg_poll(self._p)))

    def wait(self):
        _check(load_library().lg_wait(self._p))

    def set_work_unit(self, cells_per_lane: int):
        """64 (a lane steps a u64 word) or 32 (half a word, twice the lanes)."""
        _check(load_library().lg_set_work_unit(self._p, int(cells_per_lane)))

    def set_passes(self, passes: int):
        """2 (each launch collides one step and draws the next step's random
        columns), 1 (draw and collide fused per lane) or 0 (by size, the
        default)."""
        _check(load_library().lg_set_passes(self._p, int(passes)))

    def pipeline(self):
        """(kernels per step, cells per lane, draw cells per lane, draw-ahead)
        the next step will use."""
        v = [C.c_uint32() for _ in range(4)]
        _check(load_library().lg_get_pipeline(self._p, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_draw_unit(self, cells_per_lane: int):
        """64 or 32 cells per lane for the draw work of the two-pass step; 0
        (the default) by lattice size."""
        _check(load_library().lg_set_draw_unit(self._p, int(cells_per_lane)))

    def set_draw_ahead(self, steps: int):
        """0: draws share the collide launches; N: draws of up to N steps per
        launch on a second stream, ahead of the collides. Until set (or after
        set_passes(0)) it follows the lattice size."""
        _check(load_library().lg_set_draw_ahead(self._p, int(steps)))

    def set_temporal_depth(self, steps: int, rows_per_tile: int = 0):
        """Collide `steps` steps per launch on bands of rows held in registers
        (lg_set_temporal_depth); 1 turns it off."""
        _check(load_library().lg_set_temporal_depth(self._p, int(steps), int(rows_per_tile)))

    def set_resident(self, on):
        """Resident bands (lg_set_resident): one persistent launch per chunk.
        True/1 on, False/0 off, 2 by the size rule (the default)."""
        _check(load_library().lg_set_resident(self._p, int(on)))

    def resident(self):
        """(rows per band, bands) if the next advance runs resident, else None."""
        r, n = C.c_uint32(), C.c_uint32()
        on = _check(load_library().lg_get_resident(self._p, C.byref(r), C.byref(n)))
        return (r.value, n.value) if on else None

    def temporal_depth(self):
        """(steps per collide launch, valid rows per tile; 0 = by size)."""
        k, r = C.c_uint32(), C.c_uint32()
        _check(load_library().lg_get_temporal_depth(self._p, C.byref(k), C.byref(r)))
        return k.value, r.value

    def autotune(self, budget_steps: int = 0) -> int:
        """Time the step pipelines on this lattice and keep the fastest
        (lg_autotune); returns the steps taken (real steps of the lattice)."""
        used = C.c_uint64()
        _check(load_library().lg_autotune(self._p, int(budget_steps), C.byref(used)))
        return int(used.value)

    def advance_timed(self, n: int, sample_every: int = 0, alpha: float = 0.0) -> float:
        """Advance and return the device milliseconds of the step kernels,
        from HIP event pairs on the engine's stream around each run of step
        launches."""
        ms = C.c_float()
        _check(load_library().lg_advance_sampling_timed(self._p, int(n), int(sample_every),
                                                        float(alpha), C.byref(ms)))
        return ms.value

    # src/gpu.rs:710-721
    def sample_field(self, alpha: float):
        _check(load_library().lg_sample_field(self._p, float(alpha)))

    def _nblocks(self):
        if self._field is None:
            raise LgcaError("no field attached to this lattice")
        return self._field[2] * self._field[3]

    def field_dims(self):
        """(bw, bh, bx, by) of the attached field."""
        v = [C.c_size_t() for _ in range(4)]
        _check(load_library().lg_field_dims(self._p, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    # src/gpu.rs:725-736 (returns the four arrays instead of filling a Field)
    def read_field(self):
        n = self._nblocks()
        arrs = [np.zeros(n, dtype=np.float32) for _ in range(4)]
        _check(load_library().lg_read_field(self._p, *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs], n))
        return tuple(arrs)

    def read_block_sums(self):
        n = self._nblocks()
        arrs = [np.zeros(n, dtype=np.int64) for _ in range(4)]
        _check(load_library().lg_read_block_sums(self._p, *[a.ctypes.data_as(C.POINTER(C.c_int64)) for a in arrs], n))
        return tuple(arrs)

    # src/gpu.rs:740-754
    def total_particles(self) -> int:
        v = C.c_uint64()
        _check(load_library().lg_total_particles(self._p, C.byref(v)))
        return int(v.value)

    @property
    def steps(self) -> int:
        return int(load_library().lg_steps(self._p))

    @steps.setter
    def steps(self, v: int):
        _check(load_library().lg_set_steps(self._p, int(v)))


class Batch:
    """Advance several GpuLattice members together: one tile-table step kernel
    per step, the step loop replayed from a hipGraph. Members end exactly
    where their own advance() would leave them."""

    def __init__(self, members):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m._p for m in self.members])
        self._p = load_library().lg_batch_create(arr, len(self.members))
        if not self._p:
            raise LgcaError(load_library().lg_last_error().decode())

    def close(self):
        p = getattr(self, "_p", None)
        if p:
            load_library().lg_batch_destroy(p)
            self._p = None

    def __del__(self):
        self.close()

    def advance(self, n: int):
        _check(load_library().lg_batch_advance(self._p, int(n)))

    def set_shape(self, tile: int, grid: int, chunk: int):
        _check(load_library().lg_batch_set_shape(self._p, int(tile), int(grid), int(chunk)))

    def advance_limited(self, n: int, limits):
        """Member i advances min(n, limits[i]) steps."""
        _check(load_library().lg_batch_advance_limited(self._p, int(n), self._limits(limits)))

    def _limits(self, limits):
        if limits is None:
            return None
        lim = np.ascontiguousarray(limits, dtype=np.uint64)
        if lim.size != len(self.members):
            raise LgcaError("one limit per member")
        return lim.ctypes.data_as(C.POINTER(C.c_uint64))

    def advance_async(self, n: int, limits=None):
        """advance_limited (limits None: n for every member) without the wait:
        returns once the steps are queued; a member's own later calls wait
        for them first."""
        _check(load_library().lg_batch_advance_async(self._p, int(n), self._limits(limits)))

    def poll(self) -> bool:
        """True once the batch's queued work has ended."""
        return bool(_check(load_library().lg_batch_poll(self._p)))

    def wait(self):
        _check(load_library().lg_batch_wait(self._p))

    def snapshot(self, alpha: float):
        """Sample every member's attached field in one launch; returns, per
        member, (ux, uy, rho, solid) or None for a member without a field."""
        L = load_library()
        n = int(L.lg_batch_snapshot_size(self._p))
        buf = np.empty(n, dtype=np.float32)
        _check(L.lg_batch_snapshot(self._p, float(alpha), buf.ctypes.data_as(C.POINTER(C.c_float)), n))
        return self._split(buf)

    def snapshot_noalloc(self, alpha: float):
        """snapshot() through lg_batch_snapshot_noalloc: staged in the
        library's static device array, no device allocation."""
        L = load_library()
        n = int(L.lg_batch_snapshot_size(self._p))
        buf = np.empty(n, dtype=np.float32)
        _check(L.lg_batch_snapshot_noalloc(self._p, float(alpha),
                                           buf.ctypes.data_as(C.POINTER(C.c_float)), n))
        return self._split(buf)

    def advance_sampling_async(self, n: int, limits=None, alpha: float = 1.0):
        """advance_async, then the sample of snapshot() queued behind it; fetch
        the result with read_snapshot()."""
        _check(load_library().lg_batch_advance_sampling_async(self._p, int(n), self._limits(limits),
                                                              float(alpha)))

    def read_snapshot(self):
        """Wait for the batch and return the snapshot its last
        advance_sampling_async took, per member as snapshot() does."""
        L = load_library()
        n = int(L.lg_batch_snapshot_size(self._p))
        buf = np.empty(n, dtype=np.float32)
        _check(L.lg_batch_read_snapshot(self._p, buf.ctypes.data_as(C.POINTER(C.c_float)), n))
        return self._split(buf)

    def _split(self, buf):
        out, off = [], 0
        for m in self.members:
            if m._field is None:
                out.append(None)
                continue
            nb = m._field[2] * m._field[3]
            out.append(tuple(buf[off + i * nb: off + (i + 1) * nb] for i in range(4)))
            off += 4 * nb
        return out

    def set_units(self, cells_per_lane: int):
        _check(load_library().lg_batch_set_units(self._p, int(cells_per_lane)))

    def units(self) -> int:
        u = C.c_uint32()
        _check(load_library().lg_batch_get_units(self._p, C.byref(u)))
        return u.value

    def set_waves(self, waves_per_simd: int):
        """Register target of the step kernel (waves per SIMD; 0 = default)."""
        _check(load_library().lg_batch_set_waves(self._p, int(waves_per_simd)))

    def waves(self) -> int:
        w = C.c_uint32()
        _check(load_library().lg_batch_get_waves(self._p, C.byref(w)))
        return w.value

    def set_passes(self, passes: int):
        _check(load_library().lg_batch_set_passes(self._p, int(passes)))

    def set_draw_unit(self, cells_per_lane: int):
        _check(load_library().lg_batch_set_draw_unit(self._p, int(cells_per_lane)))

    def pipeline(self):
        """(kernels per step, collide cells per lane, draw cells per lane)."""
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(load_library().lg_batch_get_pipeline(self._p, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def shape(self):
        t, g, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _check(load_library().lg_batch_get_shape(self._p, C.byref(t), C.byref(g), C.byref(c)))
        return t.value, g.value, c.value

    def sample(self, alpha: float):
        _check(load_library().lg_batch_sample(self._p, float(alpha)))

    def autotune(self, budget_steps: int = 0) -> int:
        used = C.c_uint64()
        _check(load_library().lg_batch_autotune(self._p, int(budget_steps), C.byref(used)))
        return int(used.value)


def trim_memory() -> int:
    """Give the device and page-locked memory the library keeps for reuse
    back to the runtime (lg_trim_memory); returns the bytes given back."""
    v = C.c_size_t()
    _check(load_library().lg_trim_memory(C.byref(v)))
    return int(v.value)


def measure_transport(density: float, rest: bool, threads: int, seed: int, size: int = 256):
    """transport::measure (src/transport.rs:291-319) on the engine: (nu, g),
    equal to the CPU path's at the same thread count."""
    nu, g = C.c_float(), C.c_float()
    _check(load_library().lg_measure_transport(float(density), int(rest), int(threads),
                                               int(seed) & (2**64 - 1), int(size),
                                               C.byref(nu), C.byref(g)))
    return nu.value, g.value


# ---- frame output (src/render.rs:121-266, src/png.rs, src/main.rs:196-281) ----

def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _field_arrays(*arrs):
    return [np.ascontiguousarray(a, dtype=np.float32).ravel() for a in arrs]


def _path(p) -> bytes:
    return os.fsencode(p)


def field_vorticity(ux, uy, bw, bh, bx, by) -> np.ndarray:
    """Field::vorticity (src/render.rs:122-139)."""
    ux, uy = _field_arrays(ux, uy)
    out = np.zeros(bw * bh, dtype=np.float32)
    _check(load_library().lg_field_vorticity(_fptr(ux), _fptr(uy), bw, bh, bx, by, _fptr(out)))
    return out


def field_mean_velocity(ux, uy, solid):
    """Field::mean_velocity (src/render.rs:142-157)."""
    ux, uy, solid = _field_arrays(ux, uy, solid)
    mx, my = C.c_float(), C.c_float()
    _check(load_library().lg_field_mean_velocity(_fptr(ux), _fptr(uy), _fptr(solid), ux.size,
                                                 C.byref(mx), C.byref(my)))
    return mx.value, my.value


def vorticity_scale(vort, solid) -> float:
    """The colour scale of write_frame (src/main.rs:219-229)."""
    vort, solid = _field_arrays(vort, solid)
    w = C.c_float()
    _check(load_library().lg_vorticity_scale(_fptr(vort), _fptr(solid), vort.size, C.byref(w)))
    return w.value


def png_deflate(data) -> bytes:
    """png::deflate (src/png.rs:103-162)."""
    a = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))
    n = C.c_size_t()
    L = load_library()
    _check(L.lg_png_deflate(_u8(a), a.size, None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1), dtype=np.uint8)
    _check(L.lg_png_deflate(_u8(a), a.size, _u8(out), out.size, C.byref(n)))
    return out[:n.value].tobytes()


def write_png_rgb(path, rgb: np.ndarray):
    """png::write_rgb (src/png.rs:174-197) of an (h, w, 3) u8 image."""
    img = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w = img.shape[:2]
    _check(load_library().lg_write_png_rgb(_path(path), w, h, _u8(img)))


def write_vorticity(path, field, scale: int, wmax: float):
    """render::write_vorticity (src/render.rs:176-204). `field` is (ux, uy,
    solid, bw, bh, bx, by)."""
    ux, uy, solid, bw, bh, bx, by = field
    ux, uy, solid = _field_arrays(ux, uy, solid)
    _check(load_library().lg_write_vorticity(_path(path), _fptr(ux), _fptr(uy), _fptr(solid),
                                             bw, bh, bx, by, int(scale), float(wmax)))


def write_arrows(path, field, scale: float, frame):
    """render::write_arrows (src/render.rs:209-267)."""
    ux, uy, solid, bw, bh, bx, by = field
    ux, uy, solid = _field_arrays(ux, uy, solid)
    _check(load_library().lg_write_arrows(_path(path), _fptr(ux), _fptr(uy), _fptr(solid),
                                          bw, bh, bx, by, float(scale), float(frame[0]),
                                          float(frame[1])))


class Frames:
    """Frames (src/main.rs:214-281): writer threads in the library, behind a
    queue of at most `workers` frames."""

    def __init__(self, workers: int):
        self._p = load_library().lg_frames_create(int(workers))
        if not self._p:
            raise LgcaError(load_library().lg_last_error().decode())

    def write(self, field, png, svg, scale: int, arrow_scale: float, mean):
        """Frames::write (src/main.rs:262-268); `field` as for write_vorticity."""
        ux, uy, solid, bw, bh, bx, by = field
        ux, uy, solid = _field_arrays(ux, uy, solid)
        _check(load_library().lg_frames_write(self._p, _fptr(ux), _fptr(uy), _fptr(solid), bw, bh,
                                              bx, by, _path(png), _path(svg), int(scale),
                                              float(arrow_scale), float(mean[0]), float(mean[1])))

    def write_lattice(self, lattice: "GpuLattice", png, svg, scale: int, arrow_scale: float):
        """The run loop's frame (src/main.rs:676-688) from the lattice's field."""
        _check(load_library().lg_frames_write_lattice(self._p, lattice._p, _path(png), _path(svg),
                                                      int(scale), float(arrow_scale)))

    def finish(self):
        """Frames::finish (src/main.rs:271-280)."""
        p, self._p = self._p, None
        if p:
            _check(load_library().lg_frames_finish(p))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.finish()

    def __del__(self):
        if getattr(self, "_p", None):
            load_library().lg_frames_finish(self._p)
            self._p = None


def measure_viscosity(density, rest, threads, seed, size=256) -> float:
    nu = C.c_float()
    _check(load_library().lg_measure_viscosity(float(density), int(rest), int(threads),
                                               int(seed) & (2**64 - 1), int(size), C.byref(nu)))
    return nu.value


def measure_advection(density, rest, threads, seed, size=256) -> float:
    g = C.c_float()
    _check(load_library().lg_measure_advection(float(density), int(rest), int(threads),
                                               int(seed) & (2**64 - 1), int(size), C.byref(g)))
    return g.value


def collide_table_on_device(rest: bool) -> np.ndarray:
    """The generated circuit evaluated on the GPU for every (state, k)."""
    ns, stride = (128, 30) if rest else (64, 6)
    out = np.zeros(ns * stride, dtype=np.uint8)
    _check(load_library().lg_debug_collide_table(int(rest), _u8(out), out.size))
    return out.reshape(ns, stride)


PROBE_STREAM = 100  # LG_PROBE_STREAM
PROBE_OPS = 15


def probe_valu(op: int, waves_per_simd: int, chains: int = 8, iters: int = 4096):
    """The VALU issue-rate probe (lg_probe_valu): (per CU per clock, clock GHz,
    instruction name). For op PROBE_STREAM the rate is cells per CU per clock
    of the draw work's exact random stream."""
    L = load_library()
    rate, ghz = C.c_double(), C.c_double()
    _check(L.lg_probe_valu(int(op), int(waves_per_simd), int(chains), int(iters), C.byref(rate),
                           C.byref(ghz)))
    return rate.value, ghz.value, L.lg_probe_op_name(int(op)).decode()


class Sim:
    """src/sim.rs:19-175 with the GPU backend only.

    Built from the same arguments the reference's ``Sim::new`` reads off a
    prepared ``Lattice`` (src/sim.rs:60-67): size, rest flag, seed and
    thread count, the inflow, and the initial cells."""

    def __init__(self, w, h, rest, seed, threads, inlet_cols, density, inflow, cells):
        self.gpu = GpuLattice(w, h, rest, seed, threads)
        self.gpu.set_inlet(inlet_cols, density, inflow)
        self.gpu.load(cells)
        self.w, self.h = w, h
        self._field = None

    def attach_field(self, bx, by):                     # src/sim.rs:80-86
        self.gpu.attach_field(bx, by)

    def restart(self, cells, seed):                     # src/sim.rs:91-97
        self.gpu.restart(cells, seed)

    def cells(self) -> np.ndarray:                      # src/sim.rs:102-114
        return self.gpu.store()

    def advance_sampling(self, steps, sample_every, alpha):   # src/sim.rs:119-143
        self.gpu.advance_sampling(steps, max(int(sample_every), 1), alpha)
        return self.gpu.read_field()

    def sample(self, alpha):                            # src/sim.rs:148-157
        self.gpu.sample_field(alpha)
        return self.gpu.read_field()

    def advance(self, steps):                           # src/sim.rs:159-165
        self.gpu.advance(steps)

    def total_particles(self) -> int:                   # src/sim.rs:168-174
        return self.gpu.total_particles()
