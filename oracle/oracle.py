"""CPU oracle for the MI355X FFT backend -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The shipped
library (``parafaft_amd``) never imports it and has no CPU fallback.

Thin numpy/ctypes layer over ``pfx_oracle.c`` (built by ``oracle/Makefile``
into ``oracle/_build/liborc.so``).  Every function states the reference code
it restates; how each one is pinned is stated in the test that pins it
(``tests/test_oracle_*.py``).

Function                    Reference restated (file:line)
--------------------------  ---------------------------------------------------
c2c_batched                 backend/fftw3/fft_backend_fftw.hpp:387-424, 562-572
r2c_inplace / c2r_inplace   backend/fftw3/fft_backend_fftw.hpp:439-549
parafaft_c2c                parafaft_c2c.hpp:249-287, 539-694, 866-897
parafaft_r2c                parafaft_r2c.hpp:360-542, 1158-1245
decompose                   parafaft_common.hpp:73-80
chunk_range                 backend/parallel_for.hpp:53-62
stage_shapes_c2c            parafaft_c2c.hpp:781-843
stage_shapes_r2c            parafaft_r2c.hpp:1069-1151
exchange_geometry           parafaft_common.hpp:345-393
pack / unpack               parafaft_common.hpp:576-590, 654-668
subarray_serialize          parafaft_common.hpp:108-160 (MPI_ORDER_C subarray)
generate_shape              test/test_helpers.hpp:106-221
uniform                     (build-defined SplitMix64 input generator)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
_lib = None

_i64 = C.c_int64
_u64 = C.c_uint64
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_c2c_batched.argtypes = [_dp, _i64, _i64, _i64, _i64, C.c_int, C.c_int]
        L.orc_dft_brute.argtypes = [_dp, _dp, _i64, C.c_int]
        L.orc_r2c_inplace.argtypes = [_dp, _i64, _i64, _i64, _i64, C.c_int]
        L.orc_c2r_inplace.argtypes = [_dp, _i64, _i64, _i64, _i64, C.c_int]
        L.orc_decompose.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
        L.orc_chunk_range.argtypes = [_u64, C.c_int, C.c_int, _u64p, _u64p]
        L.orc_stage_shapes_c2c.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _ip]
        L.orc_stage_shapes_r2c.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip]
        L.orc_exchange_geometry.argtypes = [C.c_int, C.c_int, _ip, C.c_int, _ip, C.c_int,
                                            _ip, _ip, _ip, _ip, _u64p, _u64p, _u64p, _u64p,
                                            _u64p]
        L.orc_pack.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _ip, _ip, _u64p,
                               _u64, _u64, _u64, _u64]
        L.orc_unpack.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _ip, _ip, _u64p,
                                 _u64, _u64, _u64, _u64]
        L.orc_subarray_serialize.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _ip,
                                             C.c_int, C.c_int, C.c_int, _u64]
        L.orc_subarray_serialize.restype = C.c_uint64
        L.orc_parafaft_c2c.argtypes = [C.c_int, _ip, _dp, C.c_int, C.c_int]
        L.orc_parafaft_r2c.argtypes = [C.c_int, _ip, _dp, C.c_int, C.c_int]
        L.orc_uniform.argtypes = [_dp, _u64, _u64]
        L.orc_generate_shape.argtypes = [_dp, C.c_int, C.c_int, C.c_int]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _iarr(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.intc))
    return a, a.ctypes.data_as(_ip)


def max_threads() -> int:
    return lib().orc_max_threads()


# --------------------------------------------------------------------------
# Transforms (FP64)
# --------------------------------------------------------------------------

def c2c_batched(data: np.ndarray, length: int, batch: int, stride: int, dist: int,
                sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """In place on a complex128 buffer; element j of transform b is
    data.flat[b*dist + j*stride]."""
    assert data.dtype == np.complex128 and data.flags.c_contiguous
    rc = lib().orc_c2c_batched(_dptr(data.view(np.float64)), length, batch, stride,
                               dist, sign, nthreads)
    if rc:
        raise ValueError("orc_c2c_batched: bad arguments")
    return data


_lib_f32 = None


def lib_f32():
    """The FFT core built in single precision (oracle/_build/liborc_f32.so):
    the FFTWBackend<float> counterpart timed by the benchmark's CPU baseline."""
    global _lib_f32
    if _lib_f32 is None:
        path = os.path.join(_HERE, "_build", "liborc_f32.so")
        if not os.path.exists(path):
            build(force=True)
        L = C.CDLL(path)
        L.orc_c2c_batched.argtypes = [C.POINTER(C.c_float), _i64, _i64, _i64, _i64, C.c_int,
                                      C.c_int]
        _lib_f32 = L
    return _lib_f32


def c2c_batched_f32(data: np.ndarray, length: int, batch: int, stride: int, dist: int,
                    sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """orc_c2c_batched computed in float on complex64 data, in place."""
    assert data.dtype == np.complex64 and data.flags.c_contiguous
    lib_f32().orc_c2c_batched(data.view(np.float32).ctypes.data_as(C.POINTER(C.c_float)),
                              length, batch, stride, dist, sign, nthreads)
    return data


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """1D DFT of a contiguous vector (out of place)."""
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    return c2c_batched(y, y.size, 1, 1, y.size, sign)


def dft_brute(x: np.ndarray, sign: int = -1) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    lib().orc_dft_brute(_dptr(x.view(np.float64)), _dptr(y.view(np.float64)), x.size, sign)
    return y


def r2c_inplace(padded: np.ndarray, n: int, batch: int, stride: int, dist: int,
                nthreads: int = 0) -> np.ndarray:
    assert padded.dtype == np.float64 and padded.flags.c_contiguous
    lib().orc_r2c_inplace(_dptr(padded), n, batch, stride, dist, nthreads)
    return padded


def c2r_inplace(padded: np.ndarray, n: int, batch: int, stride: int, dist: int,
                nthreads: int = 0) -> np.ndarray:
    assert padded.dtype == np.float64 and padded.flags.c_contiguous
    lib().orc_c2r_inplace(_dptr(padded), n, batch, stride, dist, nthreads)
    return padded


def fftn(x: np.ndarray, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """D-dim unnormalised DFT through the one-rank ParaFaFT_C2C stage loop."""
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    return parafaft_c2c(y, sign, nthreads)


def parafaft_c2c(data: np.ndarray, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    assert data.dtype == np.complex128 and data.flags.c_contiguous
    shp, sp = _iarr(data.shape)
    lib().orc_parafaft_c2c(data.ndim, sp, _dptr(data.view(np.float64)), sign, nthreads)
    return data


def parafaft_r2c(padded: np.ndarray, real_shape, sign: int = -1,
                 nthreads: int = 0) -> np.ndarray:
    """padded: float64, shape real_shape[:-1] + (2*(n//2+1),), in place."""
    assert padded.dtype == np.float64 and padded.flags.c_contiguous
    shp, sp = _iarr(real_shape)
    lib().orc_parafaft_r2c(len(real_shape), sp, _dptr(padded), sign, nthreads)
    return padded


def rfftn_padded(x: np.ndarray, nthreads: int = 0) -> np.ndarray:
    """Forward R2C of a real D-dim array via the padded in-place stage loop;
    returns the complex half spectrum, shape x.shape[:-1] + (n//2+1,)."""
    n = x.shape[-1]
    h = n // 2 + 1
    pad = np.zeros(x.shape[:-1] + (2 * h,), dtype=np.float64)
    pad[..., :n] = x
    parafaft_r2c(pad, x.shape, -1, nthreads)
    return pad.view(np.complex128).reshape(x.shape[:-1] + (h,))


def irfftn_padded(X: np.ndarray, n: int, nthreads: int = 0) -> np.ndarray:
    """Backward C2R (unnormalised) of a half spectrum; returns real array."""
    h = n // 2 + 1
    assert X.shape[-1] == h
    pad = np.ascontiguousarray(X, dtype=np.complex128).copy().view(np.float64)
    pad = pad.reshape(X.shape[:-1] + (2 * h,))
    parafaft_r2c(pad, X.shape[:-1] + (n,), +1, nthreads)
    return pad[..., :n].copy()


# --------------------------------------------------------------------------
# Geometry
# --------------------------------------------------------------------------

def decompose(N: int, M: int, p: int):
    n, s = C.c_int(), C.c_int()
    lib().orc_decompose(N, M, p, C.byref(n), C.byref(s))
    return n.value, s.value


def chunk_range(total: int, tid: int, nthreads: int):
    b, e = C.c_uint64(), C.c_uint64()
    lib().orc_chunk_range(total, tid, nthreads, C.byref(b), C.byref(e))
    return b.value, e.value


def stage_shapes_c2c(global_shape, dims, coords):
    D = len(global_shape)
    g, gp = _iarr(global_shape)
    d, dp = _iarr(list(dims) + [1])
    c, cp = _iarr(list(coords) + [0])
    out = np.zeros(D * D, dtype=np.intc)
    start = np.zeros(D, dtype=np.intc)
    lib().orc_stage_shapes_c2c(D, gp, dp, cp, out.ctypes.data_as(_ip),
                               start.ctypes.data_as(_ip))
    return out.reshape(D, D).tolist(), start.tolist()


def stage_shapes_r2c(global_real, dims, coords):
    D = len(global_real)
    g, gp = _iarr(global_real)
    d, dp = _iarr(list(dims) + [1])
    c, cp = _iarr(list(coords) + [0])
    ins = np.zeros(D * D, dtype=np.intc)
    outs = np.zeros(D * D, dtype=np.intc)
    rs = np.zeros(D, dtype=np.intc)
    cs = np.zeros(D, dtype=np.intc)
    lib().orc_stage_shapes_r2c(D, gp, dp, cp, ins.ctypes.data_as(_ip),
                               outs.ctypes.data_as(_ip), rs.ctypes.data_as(_ip),
                               cs.ctypes.data_as(_ip))
    return (ins.reshape(D, D).tolist(), outs.reshape(D, D).tolist(), rs.tolist(),
            cs.tolist())


@dataclass
class ExchangeGeometry:
    """Mirror of parafaft::ExchangeGeometry (parafaft_common.hpp:273-304)."""
    src_n: list
    src_s: list
    dst_n: list
    dst_s: list
    send_counts: list
    send_displs: list
    recv_counts: list
    recv_displs: list
    src_leading: int
    src_trailing: int
    dst_leading: int
    dst_trailing: int
    src_axis_extent: int
    dst_axis_extent: int

    @property
    def src_contiguous(self):
        return self.src_leading == 1

    @property
    def dst_contiguous(self):
        return self.dst_leading == 1


def exchange_geometry(nparts, src_sizes, src_axis, dst_sizes, dst_axis) -> ExchangeGeometry:
    nd = len(src_sizes)
    s, sp = _iarr(src_sizes)
    d, dp = _iarr(dst_sizes)
    ints = [np.zeros(nparts, dtype=np.intc) for _ in range(4)]
    u = [np.zeros(nparts, dtype=np.uint64) for _ in range(4)]
    lt = np.zeros(6, dtype=np.uint64)
    lib().orc_exchange_geometry(nparts, nd, sp, src_axis, dp, dst_axis,
                                *[a.ctypes.data_as(_ip) for a in ints],
                                *[a.ctypes.data_as(_u64p) for a in u],
                                lt.ctypes.data_as(_u64p))
    return ExchangeGeometry(*[a.tolist() for a in ints], *[[int(v) for v in a] for a in u],
                            *[int(v) for v in lt])


def pack(src: np.ndarray, g: ExchangeGeometry, esz: int) -> np.ndarray:
    total = sum(g.send_counts)
    out = np.zeros(total * esz, dtype=np.uint8)
    sn, snp = _iarr(g.src_n)
    ss, ssp = _iarr(g.src_s)
    sd = np.asarray(g.send_displs, dtype=np.uint64)
    lib().orc_pack(out.ctypes.data, np.ascontiguousarray(src).ctypes.data, len(g.src_n),
                   snp, ssp, sd.ctypes.data_as(_u64p), g.src_leading, g.src_trailing,
                   g.src_axis_extent, esz)
    return out


def unpack(recv: np.ndarray, dst: np.ndarray, g: ExchangeGeometry, esz: int) -> np.ndarray:
    assert dst.flags.c_contiguous
    dn, dnp = _iarr(g.dst_n)
    ds, dsp = _iarr(g.dst_s)
    rd = np.asarray(g.recv_displs, dtype=np.uint64)
    lib().orc_unpack(dst.ctypes.data, np.ascontiguousarray(recv).ctypes.data,
                     len(g.dst_n), dnp, dsp, rd.ctypes.data_as(_u64p), g.dst_leading,
                     g.dst_trailing, g.dst_axis_extent, esz)
    return dst


def subarray_serialize(src: np.ndarray, sizes, axis: int, nparts: int, p: int,
                       esz: int) -> np.ndarray:
    n, _ = decompose(sizes[axis], nparts, p)
    total = int(np.prod(sizes)) // sizes[axis] * n
    out = np.zeros(total * esz, dtype=np.uint8)
    s, sp = _iarr(sizes)
    lib().orc_subarray_serialize(out.ctypes.data, np.ascontiguousarray(src).ctypes.data,
                                 len(sizes), sp, axis, nparts, p, esz)
    return out


# --------------------------------------------------------------------------
# Generators
# --------------------------------------------------------------------------

def uniform(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_uniform(_dptr(out), n, seed)
    return out


def uniform_complex(shape, seed: int) -> np.ndarray:
    n = int(np.prod(shape))
    return uniform(2 * n, seed).view(np.complex128).reshape(shape)


def generate_shape(D: int, N: int, shape_id: int) -> np.ndarray:
    out = np.empty(N ** D, dtype=np.float64)
    lib().orc_generate_shape(_dptr(out), D, N, shape_id)
    return out.reshape((N,) * D)


def rel_l2(y, ref) -> float:
    y = np.asarray(y, dtype=np.complex128)
    ref = np.asarray(ref, dtype=np.complex128)
    num = np.sqrt(np.sum(np.abs(y - ref) ** 2, dtype=np.float64))
    den = np.sqrt(np.sum(np.abs(ref) ** 2, dtype=np.float64))
    return float(num / den) if den > 0 else float(num)
