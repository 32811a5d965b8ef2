"""1024^3 in place on one GPU, at full size (8 GiB complex<float>, 16 GiB
complex<double>), checked through properties that hold at any size:
- the FP64 path: 8 unit plane waves exp(+2 pi i k_m . x / N) transform to
  exactly N^3 at each k_m and zero elsewhere; Parseval on iid U[-1,1) input;
  backward with the fused 1/N^3 scale returns the input;
- every FP32 form (plain, split FP16, split BF16) against that FP64 path:
  rel-L2 <= 2e-6 forward, and its own roundtrip.
The FP64 path is pinned against the oracle at 512^3 (test_gpu_config2.py);
here its full-size result is pinned by the closed forms above.
"""
import ctypes as C

import numpy as np
import pytest

from tests.gpu_util import torch

pytestmark = pytest.mark.gpu

N = 1024
SLAB = 64  # axis-0 planes per step of the norms (keeps temporaries at 1 GiB)


def _plan(prec_code):
    from parafaft_amd import _capi as A
    L = A.load()
    h = C.c_void_p()
    shape = (C.c_int * 3)(N, N, N)
    A.check(L.pfx_ndplan_create(C.byref(h), prec_code, A.PFX_C2C, 3, shape), "pfx_ndplan_create")
    return L, h


def _run(L, h, direction, x, scale=1.0):
    from parafaft_amd import _capi as A
    A.check(L.pfx_ndplan_execute(h, direction, C.c_void_p(x.data_ptr()), scale, None),
            "pfx_ndplan_execute")
    torch().cuda.synchronize()


def _sq(a):
    return float(torch().linalg.vector_norm(a.to(torch().complex128)) ** 2)


def _rel_l2(y, ref):
    """||y - ref|| / ||ref||, accumulated in FP64 over axis-0 slabs."""
    t = torch()
    num = den = 0.0
    for i in range(0, N, SLAB):
        r = ref[i:i + SLAB].to(t.complex128)
        num += float(t.linalg.vector_norm(y[i:i + SLAB].to(t.complex128) - r) ** 2)
        den += float(t.linalg.vector_norm(r) ** 2)
    return (num / den) ** 0.5


def _noise(dtype, seed):
    t = torch()
    g = t.Generator(device="cuda")
    g.manual_seed(seed)
    rdt = t.float64 if dtype == t.complex128 else t.float32
    xr = t.rand((N, N, N, 2), generator=g, device="cuda", dtype=rdt) * 2 - 1
    return t.view_as_complex(xr)


def test_fullsize_f64_closed_forms():
    from parafaft_amd import _capi as A
    t = torch()
    L, h = _plan(A.PFX_F64)
    try:
        # plane waves: exp(+2 pi i k.x / N) -> N^3 delta_k (FFTW's forward sign)
        rng = np.random.default_rng(1024)
        ks = [tuple(int(v) for v in k) for k in rng.integers(0, N, size=(8, 3))]
        g = t.arange(N, dtype=t.float64, device="cuda")
        x = t.zeros((N, N, N), dtype=t.complex128, device="cuda")
        for i in range(0, N, SLAB):
            sl = x[i:i + SLAB]
            for k in ks:
                ph = (k[0] * g[i:i + SLAB, None, None] + k[1] * g[None, :, None]
                      + k[2] * g[None, None, :])
                sl += t.exp(1j * (t.remainder(ph, N) * (2 * np.pi / N)))
        _run(L, h, A.PFX_FORWARD, x)
        peak = float(N) ** 3
        counts = {}
        for k in ks:
            counts[k] = counts.get(k, 0) + 1
        for k, c in counts.items():
            v = complex(x[k].item())
            assert abs(v - c * peak) <= 1e-9 * peak, (k, v)
            x[k] = 0
        assert float(x.abs().max()) <= 1e-9 * peak
        del x

        # Parseval and the scaled roundtrip on noise
        x = _noise(t.complex128, 7)
        x0 = x.clone()
        e_in = _sq(x)
        _run(L, h, A.PFX_FORWARD, x)
        e_out = _sq(x)
        assert abs(e_out / (peak * e_in) - 1) <= 1e-12
        _run(L, h, A.PFX_BACKWARD, x, 1.0 / peak)
        assert _rel_l2(x, x0) <= 1e-13
    finally:
        L.pfx_ndplan_destroy(h)


@pytest.mark.parametrize("mode", ["f32", "f32_split_f16", "f32_split_bf16"])
def test_fullsize_f32_forms_vs_f64(mode):
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    t = torch()
    x = _noise(t.complex64, 11)
    ref = x.to(t.complex128)
    L64, h64 = _plan(A.PFX_F64)
    try:
        _run(L64, h64, A.PFX_FORWARD, ref)
    finally:
        L64.pfx_ndplan_destroy(h64)
    L, h = _plan(P.backend.precision_code(mode))
    try:
        x0 = x.clone()
        _run(L, h, A.PFX_FORWARD, x)
        err = _rel_l2(x, ref)
        assert err <= 2e-6, (mode, err)
        del ref
        _run(L, h, A.PFX_BACKWARD, x, 1.0 / float(N) ** 3)
        # roundtrip bound 1e-6 * log2(N^3)
        assert _rel_l2(x, x0) <= 1e-6 * 30, mode
    finally:
        L.pfx_ndplan_destroy(h)
