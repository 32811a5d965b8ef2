"""Pins the oracle's transforms (CPU only).

Each oracle function is checked against (a) golden vectors from numpy's
pocketfft (tests/golden/fft_golden.npz), (b) a brute-force long-double DFT on
small sizes, and (c) closed forms of the DFT (plane waves -> exact delta
peaks, delta -> constant, Parseval, forward o backward = N I, linearity,
Hermitian symmetry of R2C).  A planted mistake (wrong sign, transposed axes,
dropped term) is shown to fail these checks.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def _rel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d else 1.0)


# -------------------------------------------------------------- brute force
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 11, 12, 16, 25, 31, 48, 64])
def test_brute_dft_matches_definition_and_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    direct = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x
    assert _rel(O.dft_brute(x), direct) < 1e-13  # numpy exp of large angles: O(n eps)
    assert _rel(O.dft_brute(x, +1), n * np.fft.ifft(x)) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 12, 13, 16, 30, 48, 60, 64, 97, 120, 128])
def test_fft_vs_brute(n):
    rng = np.random.default_rng(1000 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert _rel(O.fft(x), O.dft_brute(x)) < 1e-14
    assert _rel(O.fft(x, +1), O.dft_brute(x, +1)) < 1e-14


# -------------------------------------------------------------- golden
@pytest.mark.parametrize("name", ["c2c_8x8x8", "c2c_16x16x16", "c2c_64x64", "c2c_8x8x8x8",
                                  "c2c_4x4x4x4x4", "c2c_12x10x6"])
def test_nd_c2c_golden(name):
    x = GOLD[name + "_in"]
    assert _rel(O.fftn(x), GOLD[name + "_fwd"]) < 1e-14
    assert _rel(O.fftn(x, +1), GOLD[name + "_bwd"]) < 1e-14


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 16, 48, 80, 97, 120, 160, 1000, 1024, 1021])
def test_1d_golden(n):
    assert _rel(O.fft(GOLD[f"c2c_1d_{n}_in"]), GOLD[f"c2c_1d_{n}_fwd"]) < 1e-14


@pytest.mark.parametrize("name", ["r2c_16x12x10", "r2c_9x15", "r2c_8x8x8x8", "r2c_1d_160",
                                  "r2c_1d_81"])
def test_r2c_c2r_golden(name):
    x = GOLD[name + "_in"]
    X = O.rfftn_padded(x)
    assert _rel(X, GOLD[name + "_fwd"]) < 1e-14
    back = O.irfftn_padded(GOLD[name + "_fwd"], x.shape[-1])
    assert _rel(back, GOLD[name + "_c2r"]) < 1e-13


def test_batched_strided_geometry_matches_per_transform():
    """orc_c2c_batched with stride/dist == the 1D transform of each gathered
    sequence (FFTW guru64 semantics, fft_backend_fftw.hpp:387-424)."""
    rng = np.random.default_rng(3)
    n, batch, stride, dist = 12, 5, 7, 3
    size = (batch - 1) * dist + (n - 1) * stride + 1
    buf = rng.standard_normal(size) + 1j * rng.standard_normal(size)
    ref = buf.copy()
    for b in range(batch):
        idx = b * dist + np.arange(n) * stride
        ref[idx] = np.fft.fft(buf[idx])
    got = O.c2c_batched(buf.copy(), n, batch, stride, dist, -1)
    assert _rel(got, ref) < 1e-14


# -------------------------------------------------------------- closed forms
def test_plane_waves_give_exact_peaks():
    """sum of e^{+2 pi i k.x/N} -> N_tot at each k (config-2 construction)."""
    n = 32
    rng = np.random.default_rng(8)
    ks = rng.integers(0, n, size=(8, 3))
    g = np.indices((n, n, n)).reshape(3, -1).T
    x = np.zeros(n ** 3, np.complex128)
    for k in ks:
        x += np.exp(2j * np.pi * (g @ k) / n)
    X = O.fftn(x.reshape(n, n, n))
    expect = np.zeros((n, n, n), np.complex128)
    for k in ks:
        expect[tuple(k)] += n ** 3
    assert np.abs(X - expect).max() < 1e-9 * n ** 3


def test_delta_constant_parseval_roundtrip_linearity():
    n = (8, 6, 10)
    d = np.zeros(n, np.complex128)
    d[0, 0, 0] = 1
    assert np.abs(O.fftn(d) - 1).max() < 1e-15
    c = np.ones(n, np.complex128)
    C = O.fftn(c)
    assert abs(C[0, 0, 0] - np.prod(n)) < 1e-12 and np.abs(C.ravel()[1:]).max() < 1e-12
    x = O.uniform_complex(n, seed=4)
    y = O.uniform_complex(n, seed=5)
    X = O.fftn(x)
    assert abs(np.sum(abs(X) ** 2) - np.prod(n) * np.sum(abs(x) ** 2)) < 1e-10 * np.sum(abs(X) ** 2)
    assert _rel(O.fftn(X, +1), np.prod(n) * x) < 1e-14
    assert _rel(O.fftn(2 * x - 3j * y), 2 * X - 3j * O.fftn(y)) < 1e-14


def test_r2c_hermitian_and_truncation():
    x = O.uniform(6 * 10 * 16, seed=9).reshape(6, 10, 16)
    X = O.rfftn_padded(x)
    full = O.fftn(x.astype(np.complex128))
    assert _rel(X, full[..., :9]) < 1e-14
    # C2R o R2C = N I on real data
    assert _rel(O.irfftn_padded(X, 16), x.size * x) < 1e-14


# -------------------------------------------------------------- planted mistakes
def _closed_form_checks(transform):
    """Returns True when `transform` passes the golden + closed-form checks."""
    x = GOLD["c2c_12x10x6_in"]
    if _rel(transform(x), GOLD["c2c_12x10x6_fwd"]) > 1e-12:
        return False
    n = 16
    g = np.indices((n, n)).reshape(2, -1).T
    k = np.array([3, 5])
    w = np.exp(2j * np.pi * (g @ k) / n).reshape(n, n)
    X = transform(w)
    return abs(X[3, 5] - n * n) < 1e-9 and np.abs(X).sum() - n * n < 1e-8


def test_planted_mistakes_are_caught():
    assert _closed_form_checks(lambda x: O.fftn(x))
    # wrong sign
    assert not _closed_form_checks(lambda x: O.fftn(x, +1))
    # transposed operand: the buffer read with its shape reversed (Fortran
    # order instead of C order)
    assert not _closed_form_checks(
        lambda x: O.fftn(np.ascontiguousarray(x).reshape(x.shape[::-1])).reshape(x.shape))
    # dropped term: one axis not transformed
    def drop(x):
        y = np.ascontiguousarray(x, np.complex128).copy()
        sh = y.shape
        O.c2c_batched(y, sh[-1], y.size // sh[-1], 1, sh[-1], -1)
        return y
    assert not _closed_form_checks(drop)
    # normalised instead of FFTW's unnormalised convention
    assert not _closed_form_checks(lambda x: O.fftn(x) / x.size)


@pytest.mark.parametrize("n,batch,stride,dist", [(1024, 8, 1, 1024), (48, 20, 20, 1),
                                                  (120, 6, 1, 130), (2048, 3, 3, 1)])
def test_float_core_matches_double_oracle(n, batch, stride, dist):
    """The single-precision build of the oracle's FFT core (the benchmark's
    FP32 CPU baseline) agrees with the FP64 oracle to FP32 accuracy on the
    same inputs, both directions, in every layout."""
    size = (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size,), seed=n + batch).astype(np.complex64)
    for sign in (-1, +1):
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, stride, dist, sign)
        got = x.copy()
        O.c2c_batched_f32(got, n, batch, stride, dist, sign)
        idx = (np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride).ravel()
        assert O.rel_l2(got[idx], ref[idx]) <= 1e-6 * np.log2(n)
        # elements outside the transforms are untouched
        mask = np.ones(size, bool)
        mask[idx] = False
        assert np.array_equal(got[mask], x[mask])
