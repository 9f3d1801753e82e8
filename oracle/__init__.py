"""CPU parity oracle (test infrastructure).

ctypes binding of ``oracle/liboracle.so`` (built from ``pfdtd_oracle.c`` by ``oracle/Makefile``).
Only the tests, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
package, as the checker; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None

_p = C.c_void_p
_i64 = C.c_int64


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB_PATH))
        L.orc_to_bilbao.restype = C.c_uint8
        L.orc_to_bilbao.argtypes = [C.c_uint8, _p]
        L.orc_classify_labels.restype = None
        L.orc_classify_labels.argtypes = [_p, _i64, _i64, _i64, _p, _p, _p]
        L.orc_to_kowalczyk.restype = C.c_uint8
        L.orc_to_kowalczyk.argtypes = [C.c_uint8, _p]
        L.orc_classify_labels_kow.restype = C.c_uint64
        L.orc_classify_labels_kow.argtypes = [_p, _i64, _i64, _i64, _p, _p, _p]
        L.orc_param_vector_f.restype = None
        L.orc_param_vector_f.argtypes = [C.c_double, C.c_uint, _p]
        L.orc_param_vector_d.restype = None
        L.orc_param_vector_d.argtypes = [C.c_double, C.c_uint, _p]
        L.orc_r2y_f.restype = C.c_float
        L.orc_r2y_f.argtypes = [C.c_float]
        L.orc_dx.restype = C.c_float
        L.orc_dx.argtypes = [C.c_float, C.c_uint, C.c_double]
        L.orc_element_idx.restype = None
        L.orc_element_idx.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint, C.c_float,
                                      C.c_double, C.c_int, _p]
        L.orc_source_sample_f.restype = C.c_float
        L.orc_source_sample_f.argtypes = [C.c_int, C.c_uint, C.c_uint, _p, C.c_uint]
        L.orc_source_sample_d.restype = C.c_double
        L.orc_source_sample_d.argtypes = [C.c_int, C.c_uint, C.c_uint, _p, C.c_uint]
        L.orc_transparent_sample_f.restype = C.c_float
        L.orc_transparent_sample_f.argtypes = [C.c_int, C.c_uint, C.c_uint, _p, C.c_uint, _p, C.c_uint]
        L.orc_transparent_sample_d.restype = C.c_double
        L.orc_transparent_sample_d.argtypes = [C.c_int, C.c_uint, C.c_uint, _p, C.c_uint, _p, C.c_uint]
        L.orc_material_slot.restype = C.c_int64
        L.orc_material_slot.argtypes = [C.c_uint, C.c_uint, C.c_int]
        L.orc_run_f.restype = C.c_int
        L.orc_run_f.argtypes = [_p, _p, _i64, _i64, _i64, _p, C.c_int, C.c_uint, C.c_double,
                                _p, _p, _p, C.c_int, _p, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_int, _p, _p, _p]
        L.orc_run_d.restype = C.c_int
        L.orc_run_d.argtypes = [_p, _p, _i64, _i64, _i64, _p, C.c_int, C.c_uint, C.c_double,
                                _p, _p, _p, C.c_int, _p, C.c_int, C.c_int, C.c_int, C.c_int,
                                _p, _p, _p]
        L.orc_capture_slice_f.restype = C.c_int
        L.orc_capture_slice_f.argtypes = [_p, _p, _i64, _i64, _i64, _i64, C.c_int, _p, _p]
        L.orc_step_calls_f.restype = C.c_int
        L.orc_step_calls_f.argtypes = [_p, _p, _i64, _i64, _i64, _p, C.c_int, C.c_uint,
                                       C.c_double, _p, _p, _p, C.c_int, _p, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, _p, _p, C.c_int, _p,
                                       _p, _p]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f32_to_bf16.restype = C.c_uint16
        L.orc_f32_to_bf16.argtypes = [C.c_float]
        L.orc_bf16_to_f32.restype = C.c_float
        L.orc_bf16_to_f32.argtypes = [C.c_uint16]
        L.orc_voxel_grid.restype = C.c_int
        L.orc_voxel_grid.argtypes = [_p, _i64, C.c_float, _p, _p]
        L.orc_voxelize.restype = C.c_int
        L.orc_voxelize.argtypes = [_p, _i64, _p, _i64, _p, C.c_float, _p, _p, _p]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def voxel_grid(vertices, dx: float, pad=(1, 1, 1)) -> tuple[int, int, int]:
    """(X, Y, Z) of the voxel grid of a vertex array ([n][3] float32)."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    pad_a = np.asarray(pad, np.int32)
    dims = np.zeros(3, np.int64)
    rc = lib().orc_voxel_grid(_ptr(v), len(v), dx, _ptr(pad_a), _ptr(dims))
    if rc:
        raise ValueError(f"orc_voxel_grid failed ({rc})")
    return tuple(int(d) for d in dims)


def voxelize(vertices, indices, dx: float, tri_material=None, pad=(1, 1, 1)) -> np.ndarray:
    """Label grid (Z, Y, X) of a closed triangle mesh: 0 air (inside), m+1 solid of material m."""
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    m = None if tri_material is None else np.ascontiguousarray(tri_material, np.uint8)
    X, Y, Z = voxel_grid(v, dx, pad)
    lab = np.empty((Z, Y, X), np.uint8)
    pad_a = np.asarray(pad, np.int32)
    dims = np.zeros(3, np.int64)
    rc = lib().orc_voxelize(_ptr(v), len(v), _ptr(t), len(t), _ptr(m), dx, _ptr(pad_a),
                            _ptr(dims), _ptr(lab))
    if rc:
        raise ValueError(f"orc_voxelize failed ({rc})")
    return lab


def classify(labels: np.ndarray):
    """(pos, mat, counts[9]) for a (Z, Y, X) label grid."""
    labels = np.ascontiguousarray(labels, np.uint8)
    Z, Y, X = labels.shape
    pos = np.empty_like(labels)
    mat = np.empty_like(labels)
    counts = np.zeros(9, np.uint64)
    lib().orc_classify_labels(_ptr(labels), X, Y, Z, _ptr(pos), _ptr(mat), _ptr(counts))
    return pos, mat, counts


def classify_kow(labels: np.ndarray):
    """Centred-scheme (update type SRL) (pos, mat, counts[3]) for a (Z, Y, X) label grid."""
    labels = np.ascontiguousarray(labels, np.uint8)
    Z, Y, X = labels.shape
    pos = np.empty_like(labels)
    mat = np.empty_like(labels)
    counts = np.zeros(3, np.uint64)
    lib().orc_classify_labels_kow(_ptr(labels), X, Y, Z, _ptr(pos), _ptr(mat), _ptr(counts))
    return pos, mat, counts


STORE_F32, STORE_F16, STORE_BF16, STORE_F16_DIFF, STORE_F16_SUMDIFF = 0, 1, 2, 3, 4
SCHEME_KOWALCZYK = 4  # bit of `literal`: update type SRL (centred boundary)


def run(pos, mat, table, src_xyz, src_kind, signals, rec_xyz, n_steps, *, precision="f32",
        octave=0, lam=math.sqrt(1.0 / 3.0), literal=0, storage=STORE_F32, scale_exp=0,
        threads=1, want_fields=False):
    """Reference step loop on the CPU. Returns responses [n_rec][n_steps] (and the final
    current/past fields when want_fields)."""
    pos = np.ascontiguousarray(pos, np.uint8)
    mat = np.ascontiguousarray(mat, np.uint8)
    Z, Y, X = pos.shape
    dt = np.float64 if precision == "f64" else np.float32
    t = np.ascontiguousarray(table, np.float32)
    if t.ndim == 1:
        t = t[:, None]
    tab = np.zeros((t.shape[0], 20), np.float32)
    tab[:, : t.shape[1]] = t
    tab = np.ascontiguousarray(tab.astype(dt))
    src_xyz = np.ascontiguousarray(src_xyz, np.int32).reshape(-1, 3)
    src_kind = np.ascontiguousarray(src_kind, np.int32).reshape(-1)
    rec_xyz = np.ascontiguousarray(rec_xyz, np.int32).reshape(-1, 3)
    n_src, n_rec = src_xyz.shape[0], rec_xyz.shape[0]
    sig = np.zeros((n_src, n_steps), dt)
    if n_src:
        s = np.asarray(signals).reshape(n_src, -1)
        k = min(n_steps, s.shape[1])
        sig[:, :k] = s[:, :k].astype(dt)
    sig = np.ascontiguousarray(sig)
    resp = np.zeros((n_rec, n_steps), dt)
    P = np.zeros((Z, Y, X), dt) if want_fields else None
    Pp = np.zeros((Z, Y, X), dt) if want_fields else None
    if precision == "f64":
        rc = lib().orc_run_d(_ptr(pos), _ptr(mat), X, Y, Z, _ptr(tab), tab.shape[0], octave, lam,
                             _ptr(src_xyz), _ptr(src_kind), _ptr(sig), n_src, _ptr(rec_xyz),
                             n_rec, n_steps, literal, threads, _ptr(resp), _ptr(P), _ptr(Pp))
    else:
        rc = lib().orc_run_f(_ptr(pos), _ptr(mat), X, Y, Z, _ptr(tab), tab.shape[0], octave, lam,
                             _ptr(src_xyz), _ptr(src_kind), _ptr(sig), n_src, _ptr(rec_xyz),
                             n_rec, n_steps, literal, storage, scale_exp, threads, _ptr(resp),
                             _ptr(P), _ptr(Pp))
    if rc != 0:
        raise ValueError(f"oracle run failed with {rc}")
    return (resp, P, Pp) if want_fields else resp


def _table20(table, dt):
    t = np.ascontiguousarray(table, np.float32)
    if t.ndim == 1:
        t = t[:, None]
    tab = np.zeros((t.shape[0], 20), np.float32)
    tab[:, : t.shape[1]] = t
    return np.ascontiguousarray(tab.astype(dt))


def step_calls(pos, mat, table, src_xyz, src_kind, signals, rec_xyz, n_steps, steps, dirs, *,
               octave=0, lam=math.sqrt(1.0 / 3.0), literal=0, storage=STORE_F32, scale_exp=0,
               threads=1):
    """Single-step mode (launchFDTD3dStep per call, FP32): call c advances with step index
    steps[c] and direction dirs[c]. Returns (responses [n_rec][n_steps], current, past)."""
    pos = np.ascontiguousarray(pos, np.uint8)
    mat = np.ascontiguousarray(mat, np.uint8)
    Z, Y, X = pos.shape
    tab = _table20(table, np.float32)
    src_xyz = np.ascontiguousarray(src_xyz, np.int32).reshape(-1, 3)
    src_kind = np.ascontiguousarray(src_kind, np.int32).reshape(-1)
    rec_xyz = np.ascontiguousarray(rec_xyz, np.int32).reshape(-1, 3)
    n_src, n_rec = src_xyz.shape[0], rec_xyz.shape[0]
    sig = np.zeros((n_src, n_steps), np.float32)
    if n_src:
        s = np.asarray(signals).reshape(n_src, -1)
        k = min(n_steps, s.shape[1])
        sig[:, :k] = s[:, :k].astype(np.float32)
    st = np.ascontiguousarray(steps, np.int32)
    dr = np.ascontiguousarray(dirs, np.int32)
    assert st.shape == dr.shape
    resp = np.zeros((n_rec, n_steps), np.float32)
    P = np.zeros((Z, Y, X), np.float32)
    Pp = np.zeros((Z, Y, X), np.float32)
    rc = lib().orc_step_calls_f(_ptr(pos), _ptr(mat), X, Y, Z, _ptr(tab), tab.shape[0], octave,
                                lam, _ptr(src_xyz), _ptr(src_kind), _ptr(sig), n_src,
                                _ptr(rec_xyz), n_rec, n_steps, literal, storage, scale_exp,
                                threads, _ptr(st), _ptr(dr), st.size, _ptr(resp), _ptr(P),
                                _ptr(Pp))
    if rc != 0:
        raise ValueError(f"oracle step calls failed with {rc}")
    return resp, P, Pp


def capture_slice(P, pos, slice_, orientation):
    """orc_capture_slice_f: (p, pos) planes of an FP32 field, or None when out of range."""
    P = np.ascontiguousarray(P, np.float32)
    pos = np.ascontiguousarray(pos, np.uint8)
    Z, Y, X = P.shape
    shape = {0: (Y, X), 1: (Z, X), 2: (Z, Y)}.get(orientation, (1, 1))
    op = np.full(shape, np.nan, np.float32)
    oq = np.full(shape, 0xEE, np.uint8)
    rc = lib().orc_capture_slice_f(_ptr(P), _ptr(pos), X, Y, Z, slice_, orientation, _ptr(op),
                                   _ptr(oq))
    return None if rc else (op, oq)
