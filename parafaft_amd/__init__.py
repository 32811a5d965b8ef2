"""parafaft_amd -- MI355X (gfx950) FFT backend for ParaFaFT.

The product is ``libpfx.so`` (C ABI, include/pfx_capi.h) with hand-written
HIP kernels, and the C++ drop-in backend class in
include/parafaft_mi355/fft_backend_mi355.hpp.  This Python package is the
host-side mirror used by the tests and the benchmark: a ctypes binding of the
C ABI, the backend concept (``MI355Backend``) and one-rank drivers
(``ParaFaFT_C2C``, ``ParaFaFT_R2C``).  It has no CPU fallback: if libpfx.so is
missing, every entry point raises.
"""
from ._capi import PfxError, load as load_library, LIB_PATH  # noqa: F401
from .backend import (DeviceBuffer, FFTDirection, FFTPlanFlag, MI355Backend,  # noqa: F401
                      Plan, ptr)
from .parafaft import ParaFaFT_C2C, ParaFaFT_R2C  # noqa: F401

__all__ = ["PfxError", "load_library", "LIB_PATH", "DeviceBuffer", "FFTDirection",
           "FFTPlanFlag", "MI355Backend", "Plan", "ptr", "ParaFaFT_C2C", "ParaFaFT_R2C"]
