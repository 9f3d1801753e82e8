"""Python host object over one C-ABI context (``pfdtd_ctx``).

This is the thin host layer the tests, the benchmark and the serving loop use; the
reference-shaped ``FDTD::App`` facade (``csrc/app``, ``libpfdtd_app.so`` / ``libPyFDTD``)
calls the same C ABI from C++. Arrays cross the boundary as C-contiguous numpy buffers; all
compute happens in libpfdtd.so.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import native as N


def voxel_grid(vertices: np.ndarray, dx: float, pad=(1, 1, 1)) -> tuple[int, int, int]:
    """(X, Y, Z) of the voxel grid pfdtd_voxelize builds for these vertices."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    pad_a = np.ascontiguousarray(pad, np.int32)
    dims = np.zeros(3, np.int64)
    N.check(N.lib().pfdtd_voxel_grid(N.ptr(v), len(v), C.c_float(dx), N.ptr(pad_a), N.ptr(dims)),
            "pfdtd_voxel_grid")
    return tuple(int(d) for d in dims)


class Solver:
    """One device context: domain, materials, sources, receivers, fields, responses."""

    def __init__(self, device: int = 0):
        self._lib = N.lib()
        h = C.c_void_p()
        N.check(self._lib.pfdtd_ctx_create(device, C.byref(h)), "pfdtd_ctx_create")
        self._h = h
        self.device = device
        self.dims: tuple[int, int, int] | None = None
        self.n_rec = 0

    # -- lifetime -------------------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            N.check(self._lib.pfdtd_ctx_destroy(self._h), "pfdtd_ctx_destroy")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, name: str, *args) -> None:
        N.check(getattr(self._lib, name)(self._h, *args), name)

    # -- domain ---------------------------------------------------------------------------
    def set_domain(self, pos: np.ndarray, mat: np.ndarray) -> None:
        """Pre-classified (Z, Y, X) position and material bytes."""
        pos = np.ascontiguousarray(pos, np.uint8)
        mat = np.ascontiguousarray(mat, np.uint8)
        assert pos.shape == mat.shape and pos.ndim == 3
        Z, Y, X = pos.shape
        self._c("pfdtd_set_domain", N.ptr(pos), N.ptr(mat), X, Y, Z)
        self.dims = (X, Y, Z)

    def set_domain_labels(self, labels: np.ndarray) -> None:
        """(Z, Y, X) label grid (0 inside, m+1 solid of material m), classified on the GPU."""
        labels = np.ascontiguousarray(labels, np.uint8)
        Z, Y, X = labels.shape
        self._c("pfdtd_set_domain_labels", N.ptr(labels), X, Y, Z)
        self.dims = (X, Y, Z)

    def voxelize(self, vertices: np.ndarray, indices: np.ndarray, dx: float,
                 tri_material: np.ndarray | None = None, pad=(1, 1, 1),
                 want_labels: bool = False) -> np.ndarray | None:
        """Voxelize a closed triangle mesh (vertices [n][3] metres, indices [m][3], one material
        index per triangle) on the GPU and classify it as the domain; returns the (Z, Y, X)
        label grid when asked."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
        m = None if tri_material is None else np.ascontiguousarray(tri_material, np.uint8)
        pad_a = np.ascontiguousarray(pad, np.int32)
        X, Y, Z = voxel_grid(v, dx, pad)
        lab = np.empty((Z, Y, X), np.uint8) if want_labels else None
        dims = np.zeros(3, np.int64)
        self._c("pfdtd_voxelize", N.ptr(v), len(v), N.ptr(t), len(t),
                None if m is None else N.ptr(m), C.c_float(dx), N.ptr(pad_a), N.ptr(dims),
                None if lab is None else N.ptr(lab))
        assert tuple(int(d) for d in dims) == (X, Y, Z)
        self.dims = (X, Y, Z)
        return lab

    def set_domain_codes(self, codes: np.ndarray, mat: np.ndarray) -> None:
        """(Z, Y, X) voxelizer node codes 0..27 and material bytes (toBilbao, or toKowalczyk for
        update type SRL, on the GPU)."""
        codes = np.ascontiguousarray(codes, np.uint8)
        mat = np.ascontiguousarray(mat, np.uint8)
        Z, Y, X = codes.shape
        self._c("pfdtd_set_domain_codes", N.ptr(codes), N.ptr(mat), X, Y, Z)
        self.dims = (X, Y, Z)

    def domain(self) -> tuple[np.ndarray, np.ndarray]:
        X, Y, Z = self.dims
        pos = np.empty((Z, Y, X), np.uint8)
        mat = np.empty((Z, Y, X), np.uint8)
        self._c("pfdtd_get_domain", N.ptr(pos), N.ptr(mat))
        return pos, mat

    def counts(self) -> tuple[int, int]:
        a, b = C.c_uint64(), C.c_uint64()
        self._c("pfdtd_get_counts", C.byref(a), C.byref(b))
        return a.value, b.value

    # -- materials, sources, receivers ---------------------------------------------------
    def set_materials(self, table: np.ndarray, octave: int = 0, lam: float = math.sqrt(1.0 / 3.0),
                      literal_slot: bool = False, soft_overwrite: bool = False) -> None:
        t = np.ascontiguousarray(table, np.float32)
        if t.ndim == 1:
            t = t[:, None]
        flags = (N.MATSLOT_LITERAL if literal_slot else 0) | (N.SOFT_OVERWRITE if soft_overwrite else 0)
        self._c("pfdtd_set_materials", N.ptr(t), t.shape[0], t.shape[1], octave, lam, flags)

    def set_sources(self, xyz: np.ndarray, kind: np.ndarray, signal_f64: np.ndarray | None = None,
                    signal_f32: np.ndarray | None = None) -> None:
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        kind = np.ascontiguousarray(kind, np.int32).reshape(-1)
        n = xyz.shape[0]
        sd = None if signal_f64 is None else np.ascontiguousarray(signal_f64, np.float64).reshape(n, -1)
        sf = None if signal_f32 is None else np.ascontiguousarray(signal_f32, np.float32).reshape(n, -1)
        ns = (sd if sd is not None else sf).shape[1] if n else 0
        self._c("pfdtd_set_sources", n, N.ptr(xyz), N.ptr(kind), ns, N.ptr(sf), N.ptr(sd))

    def set_receivers(self, xyz: np.ndarray) -> None:
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        self._c("pfdtd_set_receivers", xyz.shape[0], N.ptr(xyz))
        self.n_rec = xyz.shape[0]

    # -- running --------------------------------------------------------------------------
    def prepare_fields(self, precision: int = N.F32, scale_exp: int = 0) -> None:
        self._c("pfdtd_prepare_fields", precision, scale_exp)

    def set_kernel_timing(self, on: bool) -> None:
        self._c("pfdtd_set_kernel_timing", 1 if on else 0)

    def kernel_timing(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        self._c("pfdtd_get_kernel_timing", C.byref(ms), C.byref(n))
        return ms.value, n.value

    def set_update_tiling(self, rows: int = 0, zchunk: int = 0) -> None:
        """Single-step update tiling: rows per thread and z-chunk; (0, 0) = automatic."""
        self._c("pfdtd_set_update_tiling", rows, zchunk)

    def transparent_signals(self, signals: np.ndarray, grid_ir: np.ndarray) -> np.ndarray:
        """Transparent-source filtering on the device (FP32 for float32 input, else FP64)."""
        x = np.ascontiguousarray(signals)
        f64 = x.dtype != np.float32
        x = np.ascontiguousarray(x, np.float64 if f64 else np.float32).reshape(len(x), -1)
        h = np.ascontiguousarray(grid_ir, np.float32)
        out = np.empty_like(x)
        if f64:
            self._c("pfdtd_transparent_signals", x.shape[0], x.shape[1], N.ptr(h), h.size, None,
                    None, N.ptr(x), N.ptr(out))
        else:
            self._c("pfdtd_transparent_signals", x.shape[0], x.shape[1], N.ptr(h), h.size,
                    N.ptr(x), N.ptr(out), None, None)
        return out

    def set_update_type(self, update_type: int) -> None:
        """0 SRL_FORWARD, 1 SHARED (same update), 2 SRL (centred boundaries); before the domain."""
        self._c("pfdtd_set_update_type", update_type)

    def set_batch(self, slots: int) -> None:
        self._c("pfdtd_set_batch", slots)
        self.slots = slots

    def set_batch_layout(self, layout: int) -> None:
        """N.BATCH_INTERLEAVED (slot innermost) or N.BATCH_GROUPED (slot outermost)."""
        self._c("pfdtd_set_batch_layout", layout)
        self.batch_layout = layout

    def set_source_slots(self, slots) -> None:
        sl = np.ascontiguousarray(slots, np.int32)
        self._c("pfdtd_set_source_slots", sl.size, N.ptr(sl))

    def set_receiver_slots(self, slots) -> None:
        sl = np.ascontiguousarray(slots, np.int32)
        self._c("pfdtd_set_receiver_slots", sl.size, N.ptr(sl))

    def set_time_levels(self, levels: int = 0, zchunk: int = 0) -> None:
        """Time levels per sweep (1, 2 or 3; 0 = automatic) and z-chunk (0 = automatic)."""
        self._c("pfdtd_set_time_levels", levels, zchunk)

    def set_sweep_shape(self, shape: int = N.SWEEP_AUTO) -> None:
        """Block shape of the sweep: N.SWEEP_AUTO, N.SWEEP_COLUMN, N.SWEEP_WIDE or N.SWEEP_PAIR."""
        self._c("pfdtd_set_sweep_shape", shape)

    def set_poll_interval(self, steps: int) -> None:
        self._c("pfdtd_set_poll_interval", steps)

    def run(self, n_steps: int, precision: int = N.F32, interrupt=None, progress=None) -> tuple[int, float]:
        """Run n_steps from zero fields. Returns (steps done, device milliseconds)."""
        icb = N.INTERRUPT_CB(lambda _u: int(bool(interrupt()))) if interrupt else N.INTERRUPT_CB()
        pcb = (N.PROGRESS_CB(lambda s, n, t, _u: progress(s, n, t)) if progress
               else N.PROGRESS_CB())
        done, ms = C.c_int(), C.c_double()
        self._c("pfdtd_run", precision, n_steps, icb, pcb, None, C.byref(done), C.byref(ms))
        return done.value, ms.value

    def begin_steps(self, n_steps: int, precision: int = N.F32) -> None:
        self._c("pfdtd_begin_steps", precision, n_steps)

    def step(self, step: int, direction: int = 1) -> None:
        self._c("pfdtd_step", step, direction)

    def steps(self, first_step: int, n_steps: int) -> None:
        """n_steps forward steps from first_step: the same as that many step() calls."""
        self._c("pfdtd_steps", first_step, n_steps)

    def response_shape(self) -> tuple[int, int]:
        r, s = C.c_int(), C.c_int()
        self._c("pfdtd_get_response_shape", C.byref(r), C.byref(s))
        return r.value, s.value

    def responses(self, dtype=np.float64) -> np.ndarray:
        r, s = self.response_shape()
        out = np.zeros((r, s), dtype)
        if dtype == np.float32:
            self._c("pfdtd_get_responses_f32", N.ptr(out), out.size)
        else:
            self._c("pfdtd_get_responses_f64", N.ptr(out), out.size)
        return out

    def capture_slice(self, slice_idx: int, orientation: int) -> tuple[np.ndarray, np.ndarray]:
        X, Y, Z = self.dims
        shape = {0: (Y, X), 1: (Z, X), 2: (Z, Y)}[orientation]
        p = np.zeros(shape, np.float32)
        q = np.zeros(shape, np.uint8)
        self._c("pfdtd_capture_slice", slice_idx, orientation, N.ptr(p), N.ptr(q))
        return p, q

    def capture_slice_noalloc(self, slice_idx: int, orientation: int) -> tuple[np.ndarray, np.ndarray]:
        """capture_slice staged in the library's static device buffer (no device allocation)."""
        X, Y, Z = self.dims
        shape = {0: (Y, X), 1: (Z, X), 2: (Z, Y)}[orientation]
        p = np.zeros(shape, np.float32)
        q = np.zeros(shape, np.uint8)
        self._c("pfdtd_capture_slice_noalloc", slice_idx, orientation, N.ptr(p), N.ptr(q))
        return p, q

    def capture_field(self) -> np.ndarray:
        X, Y, Z = self.dims
        out = np.zeros((Z, Y, X), np.float32)
        self._c("pfdtd_capture_field", N.ptr(out), out.size)
        return out

    def fields_raw(self, dtype) -> tuple[np.ndarray, np.ndarray]:
        X, Y, Z = self.dims
        B = getattr(self, "slots", 1)
        grouped = getattr(self, "batch_layout", N.BATCH_INTERLEAVED) == N.BATCH_GROUPED
        shape = (Z, Y, X) if B == 1 else ((B, Z, Y, X) if grouped else (Z, Y, X, B))
        p = np.zeros(shape, dtype)
        q = np.zeros(shape, dtype)
        self._c("pfdtd_get_fields_raw", N.ptr(p), N.ptr(q), p.nbytes)
        return p, q

    # -- convenience ----------------------------------------------------------------------
    def load_scene(self, scene, literal_slot: bool = False, soft_overwrite: bool = False) -> None:
        self.set_domain_labels(scene.labels)
        self.set_materials(scene.material_table(), literal_slot=literal_slot,
                           soft_overwrite=soft_overwrite)
        self.set_sources(scene.src_xyz, scene.src_kind, signal_f64=scene.signals,
                         signal_f32=scene.signals.astype(np.float32))
        self.set_receivers(scene.rec_xyz)
