"""Seeded random N-D shapes (D = 2..4, uneven extents, lengths with and
without tiled plans) through the one-rank drivers in both modes (faithful
stage loop and the fast N-D plan), C2C and padded R2C, FP32 and FP64, and
the split-precision FP16/BF16 forms on shapes with radix-16 extents:
forward against the oracle's restated stage loop, then the backward with
the fused 1/N scale back to the input."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

_EXTENTS = [2, 3, 4, 5, 6, 7, 8, 10, 12, 15, 16, 18, 20, 24, 30, 32, 36, 40, 48, 60, 64, 80,
            96, 100, 120, 128, 160, 192, 256, 384, 640, 1000]


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        D = int(rng.integers(2, 5))
        shape = [int(rng.choice(_EXTENTS)) for _ in range(D)]
        if np.prod(shape) > (1 << 21) or np.prod(shape) < 8:
            continue
        kind = "r2c" if rng.random() < 0.4 else "c2c"
        prec = "f64" if rng.random() < 0.4 else "f32"
        fast = bool(rng.random() < 0.5)
        out.append((len(out), kind, prec, fast, tuple(shape)))
    return out


@pytest.mark.parametrize("case", _cases(31337, 40), ids=lambda c: "-".join(map(str, c)))
def test_random_nd(case):
    i, kind, prec, fast, shape = case
    _run_case(i, kind, prec, fast, list(shape), tol(prec, int(np.prod(shape))))


def _run_case(i, kind, prec, fast, shape, t):
    import parafaft_amd as P
    N = int(np.prod(shape))
    base = "f64" if prec == "f64" else "f32"
    if kind == "c2c":
        cdt = np.complex64 if base == "f32" else np.complex128
        x = O.uniform_complex(tuple(shape), seed=500 + i).astype(cdt)
        ref = O.fftn(x.astype(np.complex128))
        fft = P.ParaFaFT_C2C(shape, precision=prec, fast=fast)
        d = to_dev(x)
        fft.forward(d)
        assert rel_l2(to_host(d), ref) <= t
        fft.backward(d, scale=1.0 / N)
        assert rel_l2(to_host(d), x) <= 2 * t
    else:
        rdt = np.float32 if base == "f32" else np.float64
        x = O.uniform(N, seed=600 + i).reshape(shape).astype(rdt)
        ref = O.rfftn_padded(x.astype(np.float64))
        fft = P.ParaFaFT_R2C(shape, precision=prec, fast=fast)
        h = shape[-1] // 2 + 1
        rows = int(np.prod(shape[:-1]))
        pad = np.zeros(fft.get_required_output_size(), rdt)
        pad[:rows * 2 * h].reshape(rows, 2 * h)[:, :shape[-1]] = x.reshape(rows, -1)
        d = to_dev(pad)
        fft.forward_in_place(d)
        cdt = np.complex64 if base == "f32" else np.complex128
        got = to_host(d)[:rows * 2 * h].copy().view(cdt).reshape(shape[:-1] + [h])
        assert rel_l2(got, ref) <= t
        fft.backward_in_place(d, scale=1.0 / N)
        back = to_host(d)[:rows * 2 * h].reshape(rows, 2 * h)[:, :shape[-1]]
        assert rel_l2(back.reshape(shape), x) <= 2 * t


_SPLIT_EXTENTS = [4, 12, 16, 30, 64, 100, 256, 512, 1024, 2048, 4096]


def _split_cases(seed, count):
    """Shapes with at least one extent that has radix-16 passes (the split
    forms' matrix-core kernels), the rest on the plain FP32 kernels."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        D = int(rng.integers(2, 4))
        shape = [int(rng.choice(_SPLIT_EXTENTS)) for _ in range(D)]
        if np.prod(shape) > (1 << 22) or not any(e >= 256 for e in shape):
            continue
        kind = "r2c" if rng.random() < 0.3 else "c2c"
        prec = "f32_split_bf16" if rng.random() < 0.5 else "f32_split_f16"
        fast = bool(rng.random() < 0.5)
        out.append((len(out), kind, prec, fast, tuple(shape)))
    return out


@pytest.mark.parametrize("case", _split_cases(2718, 16), ids=lambda c: "-".join(map(str, c)))
def test_random_nd_split(case):
    """The split-precision forms through both driver modes on seeded shapes:
    forward against the oracle at the FP32 bound, backward with the fused
    1/N scale back to the input."""
    i, kind, prec, fast, shape = case
    _run_case(i + 1000, kind, prec, fast, list(shape), tol("f32", int(np.prod(shape))))
