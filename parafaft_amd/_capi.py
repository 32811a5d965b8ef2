"""ctypes binding of libpfx.so (declarations in include/pfx_capi.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfx.so")

PFX_SUCCESS = 0
PFX_ERR_INVALID_ARG = 1
PFX_ERR_UNSUPPORTED = 2
PFX_ERR_HIP = 3
PFX_ERR_ALLOC = 4
PFX_ERR_INTERNAL = 5

PFX_F32 = 0
PFX_F64 = 1
PFX_F32_SPLIT_F16 = 2
PFX_F32_SPLIT_BF16 = 3

PFX_C2C = 0
PFX_R2C = 1
PFX_C2R = 2

PFX_FORWARD = 0
PFX_BACKWARD = 1

PFX_PLAN_ESTIMATE = 0
PFX_PLAN_MEASURE = 1
PFX_PLAN_PATIENT = 2

# (name, restype, argtypes) for every symbol include/pfx_capi.h declares.
_i = C.c_int
_i64 = C.c_int64
_u64 = C.c_uint64
_sz = C.c_size_t
_vp = C.c_void_p
_d = C.c_double
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)
_vpp = C.POINTER(C.c_void_p)

SIGNATURES = [
    ("pfx_error_string", C.c_char_p, [_i]),
    ("pfx_last_error_message", C.c_char_p, []),
    ("pfx_version", C.c_char_p, []),
    ("pfx_plan_create_c2c", _i, [_vpp, _i, _i64, _i64, _i64, _i64]),
    ("pfx_plan_create_c2c_many", _i, [_vpp, _i, _i64, _i64, _i64, _i64, _i64, _i64]),
    ("pfx_plan_create_r2c", _i, [_vpp, _i, _i64, _i64, _i64, _i64]),
    ("pfx_plan_create_c2r", _i, [_vpp, _i, _i64, _i64, _i64, _i64]),
    ("pfx_execute", _i, [_vp, _i, _vp, _vp]),
    ("pfx_execute_scaled", _i, [_vp, _i, _vp, _d, _vp]),
    ("pfx_execute_segmented", _i, [_vp, _i, _vp, _vp, _vp, _d, _vp]),
    ("pfx_execute_r2c_oop", _i, [_vp, _vp, _i64, _vp, _d, _vp]),
    ("pfx_execute_c2r_oop", _i, [_vp, _vp, _vp, _i64, _d, _vp]),
    ("pfx_plan_destroy", _i, [_vp]),
    ("pfx_execute_c2c_noalloc", _i, [_i, _i64, _i64, _i64, _i64, _i, _vp, _d, _vp]),
    ("pfx_plan_describe", _i, [_vp, C.c_char_p, _sz, _ip, _ip]),
    ("pfx_plan_tune", _i, [_vp, _i, _vp, _vp]),
    ("pfx_plan_num_candidates", _i, [_vp, _ip]),
    ("pfx_plan_use_candidate", _i, [_vp, _i]),
    ("pfx_tune_cache_save", _i, [C.c_char_p]),
    ("pfx_tune_cache_load", _i, [C.c_char_p]),
    ("pfx_tune_cache_dropped", _i, []),
    ("pfx_tune_cache_clear", _i, []),
    ("pfx_tune_cache_size", _i, []),
    ("pfx_malloc", _i, [_vpp, _sz]),
    ("pfx_free", _i, [_vp]),
    ("pfx_memcpy_async", _i, [_vp, _vp, _sz, _vp]),
    ("pfx_memcpy2d_async", _i, [_vp, _sz, _vp, _sz, _sz, _sz, _vp]),
    ("pfx_stream_create", _i, [_vpp]),
    ("pfx_stream_destroy", _i, [_vp]),
    ("pfx_stream_synchronize", _i, [_vp]),
    ("pfx_pack", _i, [_vp, _vp, _i, _ip, _ip, _u64p, _u64, _u64, _u64, _u64, _vp]),
    ("pfx_unpack", _i, [_vp, _vp, _i, _ip, _ip, _u64p, _u64, _u64, _u64, _u64, _vp]),
    ("pfx_scale", _i, [_vp, _i64, _i, _d, _vp]),
    ("pfx_ndplan_create", _i, [_vpp, _i, _i, _i, _ip]),
    ("pfx_ndplan_execute", _i, [_vp, _i, _vp, _d, _vp]),
    ("pfx_ndplan_execute_r2c_oop", _i, [_vp, _i, _vp, _vp, _d, _vp]),
    ("pfx_ndplan_destroy", _i, [_vp]),
    ("pfx_ndplan_stage", _i, [_vp, _i, _vpp]),
    ("pfx_ndplan_tune", _i, [_vp, _i, _vp, _vp]),
    ("pfx_ndplan_chunk", _i, [_vp, _ip]),
    ("pfx_ndplan_set_chunk", _i, [_vp, _i]),
    ("pfx_ndplan_chunk_stage", _i, [_vp, _i, _vpp]),
    ("pfx_ndplan_split", _i, [_vp, _ip]),
    ("pfx_ndplan_set_split", _i, [_vp, _i]),
    ("pfx_ndplan_split_stage", _i, [_vp, _vpp]),
    ("pfx_batch_create", _i, [_vpp, _i, _i, _vp]),
    ("pfx_batch_tune", _i, [_vp, _i, _vp]),
    ("pfx_batch_execute", _i, [_vp, _i, _vp]),
    ("pfx_batch_execute_each", _i, [_vp, _i, _vp]),
    ("pfx_batch_execute_persistent", _i, [_vp, _i, _vp]),
    ("pfx_batch_persistent_status", _i, [_vp, _vp, _ip]),
    ("pfx_batch_set_spin_limit", _i, [_vp, _i]),
    ("pfx_batch_persistent_info", _i, [_vp, _i, _ip, _ip, _ip]),
    ("pfx_batch_launches", _i, [_vp, _i, _ip, _ip]),
    ("pfx_batch_destroy", _i, [_vp]),
    ("pfx_graph_begin_capture", _i, [_vp]),
    ("pfx_graph_end_capture", _i, [_vp, _vpp]),
    ("pfx_graph_launch", _i, [_vp, _vp]),
    ("pfx_graph_destroy", _i, [_vp]),
]


class SegMap(C.Structure):
    """pfx_segmap"""
    _fields_ = [("nparts", C.c_int), ("segments", C.POINTER(C.c_void_p)),
                ("start", C.POINTER(C.c_int)), ("count", C.POINTER(C.c_int)),
                ("trailing", C.c_uint64)]


class BatchDesc(C.Structure):
    """pfx_batch_desc"""
    _fields_ = [("kind", C.c_int), ("ndim", C.c_int), ("shape", C.c_int * 4),
                ("data", C.c_void_p)]

_lib = None


class PfxError(RuntimeError):
    """Raised for any non-zero pfx status (mirrors check_hip's runtime_error,
    backend/hipfft/fft_backend_hipfft.hpp:46-51)."""

    def __init__(self, status: int, op: str, detail: str):
        super().__init__(f"pfx error in {op}: {detail} (status {status})")
        self.status = status


def load():
    """Load libpfx.so; raises if it is missing (build with
    ``python -m parafaft_amd.build``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: the HIP library must be built "
                "(python -m parafaft_amd.build); there is no CPU fallback")
        # torch ships its own HIP runtime under the same soname as the system
        # one libpfx links; whichever loads first serves the whole process, and
        # torch cannot initialise on the other.  Load torch's first so device
        # buffers from torch and libpfx's launches share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, op: str) -> None:
    if status != PFX_SUCCESS:
        L = load()
        raise PfxError(status, op, L.pfx_last_error_message().decode())
