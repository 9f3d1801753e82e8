"""ctypes binding of the C ABI in ``include/pfdtd.h`` (``libpfdtd.so``).

The library is built in-tree by ``python -m parallelfdtd_amd.build``. There is no fallback:
if the shared object is missing or a symbol is absent, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("PFDTD_LIB", Path(__file__).resolve().parent / "libpfdtd.so"))

OK = 0
E_ARG, E_HIP, E_NOMEM, E_STATE, E_DOMAIN = -1, -2, -3, -4, -5

F32, F64, F16, BF16, F16_DIFF = 0, 1, 2, 3, 4
SWEEP_AUTO, SWEEP_COLUMN, SWEEP_WIDE, SWEEP_PAIR = 0, 1, 2, 3
BATCH_INTERLEAVED, BATCH_GROUPED = 0, 1
SRC_HARD, SRC_SOFT, SRC_TRANSPARENT = 0, 1, 2
SRL_FORWARD, SHARED, SRL = 0, 1, 2  # enum UpdateType
MATSLOT_INTENDED, MATSLOT_LITERAL = 0, 1
SOFT_ADD, SOFT_OVERWRITE = 0, 2
STATIC_CAPTURE_MAX = 4194304  # PFDTD_STATIC_CAPTURE_MAX
PROBE_READ, PROBE_WRITE, PROBE_COPY, PROBE_TRIAD, PROBE_UPDATE = 0, 1, 2, 3, 4

INTERRUPT_CB = C.CFUNCTYPE(C.c_int, C.c_void_p)
PROGRESS_CB = C.CFUNCTYPE(None, C.c_int, C.c_int, C.c_float, C.c_void_p)

_p = C.c_void_p
_i32, _i64, _u32, _u64, _dbl = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
_pi32, _pi64, _pu64, _pdbl = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_uint64), C.POINTER(C.c_double)



class PfdtdJob(C.Structure):
    """struct pfdtd_job (one serving-pool job)."""
    _fields_ = [("labels", _p), ("X", _i32), ("Y", _i32), ("Z", _i32),
                ("materials", _p), ("n_mat", _i32), ("n_coef", _i32), ("octave", _u32),
                ("lambda_", _dbl), ("flags", _i32),
                ("n_src", _i32), ("src_xyz", _p), ("src_kind", _p), ("n_samples", _i32),
                ("signals", _p), ("n_rec", _i32), ("rec_xyz", _p), ("n_steps", _i32),
                ("precision", _i32), ("scale_exp", _i32)]


# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES: dict[str, list] = {
    "pfdtd_last_error": [],
    "pfdtd_version": [],
    "pfdtd_device_count": [C.POINTER(C.c_int)],
    "pfdtd_device_mem": [C.c_int, _pu64, _pu64],
    "pfdtd_ctx_create": [C.c_int, C.POINTER(_p)],
    "pfdtd_ctx_destroy": [_p],
    "pfdtd_set_domain": [_p, _p, _p, _i64, _i64, _i64],
    "pfdtd_set_domain_labels": [_p, _p, _i64, _i64, _i64],
    "pfdtd_set_domain_codes": [_p, _p, _p, _i64, _i64, _i64],
    "pfdtd_voxelize": [_p, _p, _i64, _p, _i64, _p, C.c_float, _p, _p, _p],
    "pfdtd_voxel_grid": [_p, _i64, C.c_float, _p, _p],
    "pfdtd_get_domain": [_p, _p, _p],
    "pfdtd_get_dims": [_p, _pi64, _pi64, _pi64],
    "pfdtd_get_counts": [_p, _pu64, _pu64],
    "pfdtd_set_materials": [_p, _p, C.c_int, C.c_int, C.c_uint, _dbl, C.c_int],
    "pfdtd_set_sources": [_p, C.c_int, _p, _p, C.c_int, _p, _p],
    "pfdtd_set_receivers": [_p, C.c_int, _p],
    "pfdtd_prepare_fields": [_p, C.c_int, C.c_int],
    "pfdtd_reset_fields": [_p],
    "pfdtd_run": [_p, C.c_int, C.c_int, INTERRUPT_CB, PROGRESS_CB, _p, C.POINTER(C.c_int), _pdbl],
    "pfdtd_begin_steps": [_p, C.c_int, C.c_int],
    "pfdtd_step": [_p, C.c_int, C.c_int],
    "pfdtd_steps": [_p, C.c_int, C.c_int],
    "pfdtd_get_responses_f32": [_p, _p, _i64],
    "pfdtd_get_responses_f64": [_p, _p, _i64],
    "pfdtd_get_response_shape": [_p, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pfdtd_capture_slice": [_p, _i64, C.c_int, _p, _p],
    "pfdtd_capture_slice_noalloc": [_p, _i64, C.c_int, _p, _p],
    "pfdtd_capture_field": [_p, _p, _i64],
    "pfdtd_get_fields_raw": [_p, _p, _p, _i64],
    "pfdtd_set_kernel_timing": [_p, C.c_int],
    "pfdtd_get_kernel_timing": [_p, _pdbl, _pi64],
    "pfdtd_set_poll_interval": [_p, C.c_int],
    "pfdtd_set_update_type": [_p, C.c_int],
    "pfdtd_transparent_signals": [_p, C.c_int, C.c_int, _p, C.c_int, _p, _p, _p, _p],
    "pfdtd_set_batch": [_p, C.c_int],
    "pfdtd_set_batch_layout": [_p, C.c_int],
    "pfdtd_set_source_slots": [_p, C.c_int, _p],
    "pfdtd_set_receiver_slots": [_p, C.c_int, _p],
    "pfdtd_set_update_tiling": [_p, C.c_int, C.c_int],
    "pfdtd_set_time_levels": [_p, C.c_int, C.c_int],
    "pfdtd_set_sweep_shape": [_p, C.c_int],
    "pfdtd_hbm_probe": [C.c_int, C.c_int, _i64, C.c_int, _pdbl, _pdbl],
    "pfdtd_pool_create": [C.c_int, _i64, C.POINTER(_p)],
    "pfdtd_pool_destroy": [_p],
    "pfdtd_pool_set_options": [_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "pfdtd_pool_job_bytes": [C.POINTER(PfdtdJob), _pi64],
    "pfdtd_pool_mem": [_p, _pi64, _pi64, _pi64],
    "pfdtd_pool_submit": [_p, C.POINTER(PfdtdJob), _pi64],
    "pfdtd_pool_advance": [_p, C.POINTER(C.c_int)],
    "pfdtd_pool_poll": [_p, _pi64, C.c_int, C.POINTER(C.c_int)],
    "pfdtd_pool_drain": [_p],
    "pfdtd_pool_job_info": [_p, _i64, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "pfdtd_pool_get_responses": [_p, _i64, _p, _p, _i64],
    "pfdtd_pool_release": [_p, _i64],
    "pfdtd_pool_stats": [_p, _pi64, _pi64, _pdbl, C.POINTER(C.c_int)],
}
_RESTYPE = {"pfdtd_last_error": C.c_char_p}

_lib = None


class PfdtdError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} -> {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libpfdtd.so (raises if it is missing: the HIP path has no substitute)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing; build it with `python -m parallelfdtd_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def check(rc: int, where: str) -> None:
    if rc != OK:
        raise PfdtdError(rc, where, lib().pfdtd_last_error().decode(errors="replace"))


def ptr(a: np.ndarray | None) -> C.c_void_p | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays crossing the C ABI must be C-contiguous"
    return C.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().pfdtd_device_count(C.byref(n)), "pfdtd_device_count")
    return n.value


def device_mem(device: int = 0) -> tuple[int, int]:
    f, t = C.c_uint64(0), C.c_uint64(0)
    check(lib().pfdtd_device_mem(device, C.byref(f), C.byref(t)), "pfdtd_device_mem")
    return f.value, t.value
