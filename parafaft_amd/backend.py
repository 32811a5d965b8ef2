"""Host-side mirror of the reference's FFT backend concept over libpfx.so.

``MI355Backend`` has the member set that ParaFaFT_C2C / ParaFaFT_R2C call on a
backend (backend/fft_backend.hpp:45-196 plus the undocumented members the
drivers use: memcpy2d, sync, first_touch, use_alltoallw, handles_distributed,
use_p2p), with the argument meaning and error behaviour of
``HipFFTBackend<T>`` (backend/hipfft/fft_backend_hipfft.hpp:178-635).  The
C++ drop-in with the same surface is include/parafaft_mi355/fft_backend_mi355.hpp.

Pointers are device addresses: an ``int``, or any object with ``data_ptr()``
(a torch tensor is convenient plumbing for device memory in tests).
"""
from __future__ import annotations

import ctypes as C
import enum

from . import _capi as A


class FFTDirection(enum.Enum):
    """backend/fft_backend.hpp:25-28"""
    Forward = 0
    Backward = 1


class FFTPlanFlag(enum.Enum):
    """backend/fft_backend.hpp:39-43.  Estimate: static kernel table; Measure /
    Patient: the candidate kernels are timed on the planning buffer
    (pfx_plan_tune) and the fastest is kept and cached."""
    Estimate = 0
    Measure = 1
    Patient = 2


def ptr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    if isinstance(x, C.c_void_p):
        return int(x.value or 0)
    raise TypeError(f"cannot take a device address of {type(x)!r}")


_PREC = {"f32": A.PFX_F32, "float": A.PFX_F32, "f64": A.PFX_F64, "double": A.PFX_F64,
         "f32_split_f16": A.PFX_F32_SPLIT_F16, "f32_split_bf16": A.PFX_F32_SPLIT_BF16}


def tune_cache_save(path: str):
    A.check(A.load().pfx_tune_cache_save(path.encode()), "pfx_tune_cache_save")


def tune_cache_load(path: str) -> int:
    """Loads a tuning cache file; returns the number of lines dropped (a key
    not in the cache's format, or a kernel this build does not have)."""
    A.check(A.load().pfx_tune_cache_load(path.encode()), "pfx_tune_cache_load")
    return A.load().pfx_tune_cache_dropped()


def tune_cache_clear():
    A.load().pfx_tune_cache_clear()


def tune_cache_size() -> int:
    return A.load().pfx_tune_cache_size()


def precision_code(p) -> int:
    if isinstance(p, int):
        return p
    return _PREC[str(p)]


class Plan:
    """One batched strided 1D plan (owns its pfx_plan)."""

    def __init__(self, handle: int):
        self.handle = handle

    @classmethod
    def c2c(cls, precision, length, batch, stride, dist, outer_batch=1, outer_dist=0):
        L = A.load()
        h = C.c_void_p()
        if outer_batch == 1 and outer_dist == 0:
            A.check(L.pfx_plan_create_c2c(C.byref(h), precision_code(precision), length, batch,
                                          stride, dist), "pfx_plan_create_c2c")
        else:
            A.check(L.pfx_plan_create_c2c_many(C.byref(h), precision_code(precision), length,
                                               batch, stride, dist, outer_batch, outer_dist),
                    "pfx_plan_create_c2c_many")
        return cls(h.value)

    @classmethod
    def r2c(cls, precision, length, batch, stride, dist):
        L = A.load()
        h = C.c_void_p()
        A.check(L.pfx_plan_create_r2c(C.byref(h), precision_code(precision), length, batch,
                                      stride, dist), "pfx_plan_create_r2c")
        return cls(h.value)

    @classmethod
    def c2r(cls, precision, length, batch, stride, dist):
        L = A.load()
        h = C.c_void_p()
        A.check(L.pfx_plan_create_c2r(C.byref(h), precision_code(precision), length, batch,
                                      stride, dist), "pfx_plan_create_c2r")
        return cls(h.value)

    def execute(self, direction: int, data, stream=None, scale: float = 1.0):
        L = A.load()
        if scale == 1.0:
            A.check(L.pfx_execute(self.handle, direction, ptr(data), ptr(stream)), "pfx_execute")
        else:
            A.check(L.pfx_execute_scaled(self.handle, direction, ptr(data), scale, ptr(stream)),
                    "pfx_execute_scaled")

    def tune(self, flag: int, scratch, stream=None):
        """Time the candidate kernels on `scratch` (destroyed) and keep the
        fastest (FFTW_MEASURE / FFTW_PATIENT semantics)."""
        A.check(A.load().pfx_plan_tune(self.handle, int(flag), ptr(scratch), ptr(stream)),
                "pfx_plan_tune")

    def num_candidates(self) -> int:
        n = C.c_int()
        A.check(A.load().pfx_plan_num_candidates(self.handle, C.byref(n)),
                "pfx_plan_num_candidates")
        return n.value

    def use_candidate(self, index: int):
        A.check(A.load().pfx_plan_use_candidate(self.handle, index), "pfx_plan_use_candidate")

    def describe(self):
        L = A.load()
        buf = C.create_string_buffer(256)
        pb, th = C.c_int(), C.c_int()
        A.check(L.pfx_plan_describe(self.handle, buf, 256, C.byref(pb), C.byref(th)),
                "pfx_plan_describe")
        return buf.value.decode(), pb.value, th.value

    def destroy(self):
        if self.handle:
            A.load().pfx_plan_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class MI355Backend:
    """Backend for ParaFaFT on one MI355X; see module docstring."""

    use_alltoallw = False        # drivers pack on the device (c2c:219-221, 576-580)
    handles_distributed = False  # the per-stage driver loop stays in charge
    use_p2p = True               # P2P statics present; no peer reports access

    def __init__(self, num_stages: int, plan_flag: FFTPlanFlag = FFTPlanFlag.Estimate,
                 precision="f64", stream=None):
        self.num_stages = num_stages
        self.plan_flag = plan_flag
        self.precision = precision_code(precision)
        self._plans = [None] * num_stages
        self._r2c = None
        self._c2r = None
        L = A.load()
        if stream is None:
            h = C.c_void_p()
            A.check(L.pfx_stream_create(C.byref(h)), "pfx_stream_create")
            self.stream = h.value
            self._own_stream = True
        else:
            self.stream = ptr(stream)
            self._own_stream = False

    # -- C2C ---------------------------------------------------------------
    def create_stage_plan(self, stage, length, batch, data, stride, dist):
        """fft_backend_hipfft.hpp:294-314 (data is a layout reference only)."""
        old = self._plans[stage]
        self._plans[stage] = Plan.c2c(self.precision, length, batch, stride, dist)
        self._tune(self._plans[stage], data)
        if old is not None:
            old.destroy()

    def execute_stage(self, stage, direction: FFTDirection, data, scale: float = 1.0):
        """fft_backend_hipfft.hpp:326-339: in place, async on the stream."""
        d = A.PFX_FORWARD if direction == FFTDirection.Forward else A.PFX_BACKWARD
        self._plans[stage].execute(d, data, self.stream, scale)

    # -- R2C / C2R ------------------------------------------------------------
    def _tune(self, plan, data):
        if self.plan_flag != FFTPlanFlag.Estimate and ptr(data):
            plan.tune(self.plan_flag.value, data, self.stream)

    def create_r2c_inplace_plan(self, length, batch, padded_real, stride, dist):
        self._r2c = Plan.r2c(self.precision, length, batch, stride, dist)
        self._tune(self._r2c, padded_real)

    def execute_r2c_inplace(self, padded_real, scale: float = 1.0):
        self._r2c.execute(A.PFX_FORWARD, padded_real, self.stream, scale)

    def create_c2r_inplace_plan(self, length, batch, padded_real, stride, dist):
        self._c2r = Plan.c2r(self.precision, length, batch, stride, dist)
        self._tune(self._c2r, padded_real)

    def execute_c2r_inplace(self, padded_real, scale: float = 1.0):
        self._c2r.execute(A.PFX_BACKWARD, padded_real, self.stream, scale)

    # -- memory ---------------------------------------------------------------
    def memcpy(self, dest, src, nbytes):
        A.check(A.load().pfx_memcpy_async(ptr(dest), ptr(src), nbytes, self.stream),
                "pfx_memcpy_async")

    def memcpy2d(self, dst, dpitch, src, spitch, width, height):
        A.check(A.load().pfx_memcpy2d_async(ptr(dst), dpitch, ptr(src), spitch, width, height,
                                            self.stream), "pfx_memcpy2d_async")

    def sync(self):
        A.check(A.load().pfx_stream_synchronize(self.stream), "pfx_stream_synchronize")

    def first_touch(self, data, nbytes):
        """No-op on device memory (fft_backend_hipfft.hpp:516)."""

    def close(self):
        for p in self._plans:
            if p is not None:
                p.destroy()
        self._plans = [None] * self.num_stages
        for p in (self._r2c, self._c2r):
            if p is not None:
                p.destroy()
        self._r2c = self._c2r = None
        if self._own_stream and self.stream:
            A.load().pfx_stream_destroy(self.stream)
            self.stream = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Owning device allocation (hipvector<T>, fft_backend_hipfft.hpp:58-121)."""

    def __init__(self, nbytes: int = 0):
        self.ptr = 0
        self.nbytes = 0
        if nbytes:
            self.resize(nbytes)

    def resize(self, nbytes: int):
        if nbytes == self.nbytes:
            return
        L = A.load()
        h = C.c_void_p()
        A.check(L.pfx_malloc(C.byref(h), nbytes), "pfx_malloc")
        if self.ptr:
            L.pfx_free(self.ptr)
        self.ptr = h.value or 0
        self.nbytes = nbytes

    def data_ptr(self) -> int:
        return self.ptr

    def __del__(self):
        try:
            if self.ptr:
                A.load().pfx_free(self.ptr)
                self.ptr = 0
        except Exception:
            pass
