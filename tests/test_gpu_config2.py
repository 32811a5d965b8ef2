"""512^3 in-place C2C on one rank, complex<double> and complex<float>.
Inputs: iid U[-1,1) re/im plus 8 unit plane waves exp(+2 pi i k_m . x / 512)
with random integer wavevectors.  Checks:
- the full forward against the FP64 oracle (restated FFTWBackend<double>
  stage loop): rel-L2 <= 1e-13 (FP64), <= 1e-6 log2(N) (FP32);
- the plane waves alone: exact peaks 512^3 at each k_m, zero elsewhere;
- linearity: FFT(noise + waves) = FFT(noise) + FFT(waves);
- backward with the fused 1/N^3 scale returns the input.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, torch

pytestmark = pytest.mark.gpu

N = 512


def _waves(ks, dtype):
    t = torch()
    g = t.arange(N, dtype=t.float64, device="cuda")
    out = t.zeros((N, N, N), dtype=t.complex128, device="cuda")
    for k in ks:
        ph = (k[0] * g[:, None, None] + k[1] * g[None, :, None] + k[2] * g[None, None, :])
        ph = t.remainder(ph, N) * (2 * np.pi / N)
        out += t.exp(1j * ph)
    return out.to(dtype)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_config2_512cube(prec):
    import parafaft_amd as P
    t = torch()
    cdt_t = t.complex128 if prec == "f64" else t.complex64
    cdt = np.complex128 if prec == "f64" else np.complex64
    rng = np.random.default_rng(2512)
    ks = rng.integers(0, N, size=(8, 3))
    noise = O.uniform_complex((N, N, N), seed=512).astype(cdt)
    waves = _waves(ks, cdt_t)
    fft = P.ParaFaFT_C2C([N] * 3, precision=prec, fast=True)
    # plane waves alone: exact peaks
    dw = waves.clone()
    fft.forward(dw)
    Fw = dw
    peak = float(N ** 3)
    expect = t.zeros_like(Fw)
    for k in ks:
        expect[int(k[0]), int(k[1]), int(k[2])] += peak
    err_w = float((Fw - expect).abs().max()) / peak
    assert err_w < (1e-12 if prec == "f64" else 2e-6), err_w
    # noise + waves vs the oracle, and linearity
    dn = to_dev(noise)
    fft.forward(dn)
    x = t.from_numpy(noise).to("cuda") + waves
    fft_faithful = P.ParaFaFT_C2C([N] * 3, precision=prec, fast=False)
    dx = x.clone()
    fft_faithful.forward(dx)
    lin = float(t.linalg.vector_norm(dx - (dn + Fw)) / t.linalg.vector_norm(dx))
    assert lin < (1e-14 if prec == "f64" else 1e-6), lin
    ref = O.fftn(x.cpu().numpy().astype(np.complex128))
    err = rel_l2(to_host(dx), ref)
    assert err <= (1e-13 if prec == "f64" else 1e-6 * 27), err
    # backward with the fused 1/N^3 scale
    fft.backward(dx, scale=1.0 / N ** 3)
    back = float(t.linalg.vector_norm(dx - x) / t.linalg.vector_norm(x))
    assert back <= (1e-13 if prec == "f64" else 1e-6 * 27), back
