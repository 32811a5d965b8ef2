"""GPU parity of whole transforms through the one-rank ParaFaFT drivers.

Mirrors the reference's integration matrix (test/CMakeLists.txt:215-271):
shape in {Gaussian sigma=4, step sphere, LCG polynomial} x {D2 N32, D3 N32,
D4 N16, D5 N8}, forward compared with a serial reference transform
(test/c2c/test_fftw_c2c_nd.cpp, test/r2c/test_fftw_r2c_padded_nd.cpp) and
roundtrips scaled by 1/N^D (test/c2c/test_fftw_c2c_roundtrip_nd.cpp:73-93).
The serial reference is the oracle's FFTWBackend<double> stage loop; both the
faithful driver (exchange copies, per-slice middle-axis launches) and the
one-GPU fast path are checked.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

MATRIX = [(2, 32), (3, 32), (4, 16), (5, 8)]
SHAPES = [0, 1, 2]  # Gaussian, step, polynomial (test_helpers.hpp:106-221)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("D,N", MATRIX)
@pytest.mark.parametrize("shape_id", SHAPES)
def test_c2c_nd_matrix(prec, fast, D, N, shape_id):
    import parafaft_amd as P
    x = O.generate_shape(D, N, shape_id).astype(np.complex128)
    ref = O.fftn(x)
    cdt = np.complex64 if prec == "f32" else np.complex128
    fft = P.ParaFaFT_C2C([N] * D, precision=prec, fast=fast)
    assert fft.get_local_shape() == [N] * D
    assert fft.get_required_output_size() == N ** D
    # one rank of a (D-1)-dimensional grid (parafaft_c2c.hpp:439-477)
    assert fft.get_domain_decomposition() == [1] * D
    assert fft.get_grid_ndims() == D - 1 and fft.get_grid_coords() == [0] * (D - 1)
    assert fft.get_buffer() is None
    d = to_dev(x.astype(cdt))
    fft.forward(d)
    y = to_host(d)
    # reference tolerances: max-abs < 1e-10 (double), < 1e-4 * peak (float)
    if prec == "f64":
        assert np.abs(y - ref).max() < 1e-10
    else:
        assert np.abs(y - ref).max() < 1e-4 * np.abs(ref).max()
    assert rel_l2(y, ref) <= tol(prec, N ** D) or np.abs(ref).max() == 0
    fft.backward(d, scale=1.0 / N ** D)
    back = to_host(d)
    if prec == "f64":
        assert np.abs(back - x).max() < 1e-10
    else:
        assert np.abs(back - x).max() < 1e-4 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("fast", [False, True])
@pytest.mark.parametrize("D,N", MATRIX)
@pytest.mark.parametrize("shape_id", SHAPES)
def test_r2c_nd_matrix(prec, fast, D, N, shape_id):
    import parafaft_amd as P
    x = O.generate_shape(D, N, shape_id)
    ref = O.rfftn_padded(x)
    rdt = np.float32 if prec == "f32" else np.float64
    fft = P.ParaFaFT_R2C([N] * D, precision=prec, fast=fast)
    h = N // 2 + 1
    assert fft.get_required_output_size() == N ** (D - 1) * 2 * h
    pad = np.zeros(x.shape[:-1] + (2 * h,), dtype=rdt)
    pad[..., :N] = x
    d = to_dev(pad)
    fft.forward_in_place(d)
    got = to_host(d).view(np.complex64 if prec == "f32" else np.complex128)
    if prec == "f64":
        assert np.abs(got - ref).max() < 1e-10
    else:
        assert np.abs(got - ref).max() < 1e-4 * np.abs(ref).max()
    fft.backward_in_place(d, scale=1.0 / N ** D)
    back = to_host(d)[..., :N]
    if prec == "f64":
        assert np.abs(back - x).max() < 1e-10
    else:
        assert np.abs(back - x).max() < 1e-4 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_config1_uniform_64cube(prec):
    """64^3 complex input, iid U[-1,1) re/im from a fixed seed; forward, then
    backward * 1/64^3, checked against the input and a direct transform."""
    import parafaft_amd as P
    n = 64
    x = O.uniform_complex((n, n, n), seed=20240601)
    cdt = np.complex64 if prec == "f32" else np.complex128
    xin = x.astype(cdt)
    ref = O.fftn(xin.astype(np.complex128))
    for fast in (False, True):
        fft = P.ParaFaFT_C2C([n] * 3, precision=prec, fast=fast)
        d = to_dev(xin)
        fft.forward(d)
        y = to_host(d)
        assert rel_l2(y, ref) <= tol(prec, n ** 3), rel_l2(y, ref)
        fft.backward(d, scale=1.0 / n ** 3)
        assert rel_l2(to_host(d), xin) <= tol(prec, n ** 3)


@pytest.mark.parametrize("shape", [(12, 20, 10), (7, 9, 16), (48, 6, 80)])
def test_nonpow2_c2c_and_r2c(shape):
    import parafaft_amd as P
    x = O.uniform_complex(shape, seed=sum(shape))
    for fast in (False, True):
        fft = P.ParaFaFT_C2C(list(shape), precision="f64", fast=fast)
        d = to_dev(x)
        fft.forward(d)
        assert rel_l2(to_host(d), O.fftn(x)) <= 1e-13
    r = x.real.copy()
    h = shape[-1] // 2 + 1
    for fast in (False, True):
        fft = P.ParaFaFT_R2C(list(shape), precision="f64", fast=fast)
        pad = np.zeros(shape[:-1] + (2 * h,))
        pad[..., :shape[-1]] = r
        d = to_dev(pad)
        fft.forward_in_place(d)
        got = to_host(d).view(np.complex128)
        assert rel_l2(got, O.rfftn_padded(r)) <= 1e-13


@pytest.mark.parametrize("kind", ["c2c", "r2c"])
def test_fast_measure_planning(kind, monkeypatch):
    """Fast mode with FFTPlanFlag.Measure: kernels and the chunked or
    stage-by-stage schedule are timed on a temporary buffer at construction
    (a 1 MB chunk budget brings the chunked schedule into the comparison);
    the user's array is untouched until forward, and the result matches the
    oracle."""
    import parafaft_amd as P
    from parafaft_amd.backend import FFTPlanFlag
    monkeypatch.setenv("PFX_ND_CHUNK_MB", "1")
    shape = [64, 48, 40]
    if kind == "c2c":
        x = O.uniform_complex(tuple(shape), seed=808)
        d = to_dev(x)
        fft = P.ParaFaFT_C2C(shape, precision="f64", plan_flag=FFTPlanFlag.Measure, fast=True)
        assert np.array_equal(to_host(d), x)
        fft.forward(d)
        assert rel_l2(to_host(d), O.fftn(x)) <= 1e-13
    else:
        x = O.uniform(int(np.prod(shape)), seed=809).reshape(shape)
        fft = P.ParaFaFT_R2C(shape, precision="f64", plan_flag=FFTPlanFlag.Measure, fast=True)
        h = shape[-1] // 2 + 1
        pad = np.zeros(fft.get_required_output_size())
        pad.reshape(-1)[:np.prod(shape[:-1]) * 2 * h].reshape(shape[:-1] + [2 * h])[..., :shape[-1]] = x
        d = to_dev(pad)
        fft.forward_in_place(d)
        got = to_host(d)[:np.prod(shape[:-1]) * 2 * h].view(np.complex128).reshape(shape[:-1] + [h])
        assert rel_l2(got, O.rfftn_padded(x)) <= 1e-13
