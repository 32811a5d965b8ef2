"""Out-of-place R2C / C2R: ParaFaFT_R2C::forward(real, complex) and
::backward(complex, real) (parafaft_r2c.hpp:333-339, 446-453), whose padding
copies (copy_real_to_padded / copy_padded_to_real, :804-847) the fast path
fuses into the first R2C and the last C2R pass.  Every form must equal, bit
for bit, the copy followed by the in-place transform (same kernels), match
the oracle, and leave the real input untouched."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

SHAPES = [(32, 48, 64), (64, 96), (8, 6, 10, 16), (16, 12, 30), (12, 10, 35)]  # last: odd axis


def _padded(x):
    n = x.shape[-1]
    pad = np.zeros(x.shape[:-1] + (2 * (n // 2 + 1),), x.dtype)
    pad[..., :n] = x
    return pad


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_r2c_out_of_place(prec, fast, shape):
    import parafaft_amd as P
    rdt = np.float32 if prec == "f32" else np.float64
    x = O.uniform(int(np.prod(shape)), seed=len(shape) * 100 + shape[-1]).reshape(shape).astype(rdt)
    fft = P.ParaFaFT_R2C(list(shape), precision=prec, fast=fast)
    need = fft.get_required_output_size()
    # reference: copy into padded rows, then the in-place transform
    pad = np.zeros(need, rdt)
    pad[:_padded(x).size] = _padded(x).reshape(-1)
    d_ref = to_dev(pad)
    fft.forward_in_place(d_ref)
    want = to_host(d_ref)
    # out of place
    d_real = to_dev(x)
    d_out = to_dev(np.full(need, np.nan, rdt))
    fft.forward(d_real, d_out)
    got = to_host(d_out)
    h = shape[-1] // 2 + 1
    nc = int(np.prod(shape[:-1])) * h
    assert np.array_equal(got[:2 * nc], want[:2 * nc])
    assert np.array_equal(to_host(d_real), x)          # the input is left alone
    spec = got[:2 * nc].view(np.complex64 if prec == "f32" else np.complex128)
    ref = O.rfftn_padded(x.astype(np.float64))
    assert rel_l2(spec.reshape(ref.shape), ref) <= tol(prec, int(np.prod(shape)))
    # backward: out of place vs in place + unpad, and the round trip
    N = int(np.prod(shape))
    d_in = to_dev(got.copy())
    d_back = to_dev(np.full(int(np.prod(shape)), np.nan, rdt))
    fft.backward(d_in, d_back, scale=1.0 / N)
    back = to_host(d_back).reshape(shape)
    d_ref2 = to_dev(got.copy())
    fft.backward_in_place(d_ref2, scale=1.0 / N)
    want_back = to_host(d_ref2)[:_padded(x).size].reshape(_padded(x).shape)[..., :shape[-1]]
    assert np.array_equal(back, want_back)
    assert rel_l2(back, x) <= tol(prec, N)


def test_r2c_out_of_place_chunked(monkeypatch):
    """The fused forms inside the chunked N-D schedule (chunk budget forced
    small): real rows of each chunk at their own offsets."""
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    monkeypatch.setenv("PFX_ND_CHUNK_MB", "1")
    shape = (128, 64, 96)
    x = O.uniform(int(np.prod(shape)), seed=9).reshape(shape).astype(np.float32)
    fft = P.ParaFaFT_R2C(list(shape), precision="f32", fast=True)
    chunk = C.c_int()
    A.check(A.load().pfx_ndplan_chunk(fft._nd.handle, C.byref(chunk)), "chunk")
    assert chunk.value > 0
    need = fft.get_required_output_size()
    pad = np.zeros(need, np.float32)
    pad[:_padded(x).size] = _padded(x).reshape(-1)
    d_ref = to_dev(pad)
    fft.forward_in_place(d_ref)
    d_out = to_dev(np.zeros(need, np.float32))
    fft.forward(to_dev(x), d_out)
    assert np.array_equal(to_host(d_out), to_host(d_ref))
    d_back = to_dev(np.zeros(x.size, np.float32))
    fft.backward(d_out, d_back, scale=1.0 / x.size)
    fft.backward_in_place(d_ref, scale=1.0 / x.size)
    want = to_host(d_ref)[:_padded(x).size].reshape(_padded(x).shape)[..., :shape[-1]]
    assert np.array_equal(to_host(d_back).reshape(shape), want)


def test_plan_out_of_place_entry_points():
    """pfx_execute_r2c_oop / pfx_execute_c2r_oop on single plans, and their
    argument checks."""
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    L = A.load()
    n, batch, pd = 256, 37, 2 * (256 // 2 + 1)
    x = O.uniform(n * batch, seed=5).reshape(batch, n).astype(np.float32)
    r2c = P.Plan.r2c("f32", n, batch, 1, pd)
    c2r = P.Plan.c2r("f32", n, batch, 1, pd)
    d_ref = to_dev(_padded(x))
    r2c.execute(A.PFX_FORWARD, d_ref)
    d_real = to_dev(x)
    d_pad = to_dev(np.zeros((batch, pd), np.float32))
    A.check(L.pfx_execute_r2c_oop(r2c.handle, d_real.data_ptr(), n, d_pad.data_ptr(), 1.0, None),
            "r2c_oop")
    assert np.array_equal(to_host(d_pad), to_host(d_ref))
    d_out = to_dev(np.zeros((batch, n), np.float32))
    keep = to_host(d_pad).copy()
    A.check(L.pfx_execute_c2r_oop(c2r.handle, d_pad.data_ptr(), d_out.data_ptr(), n, 0.5, None),
            "c2r_oop")
    assert np.array_equal(to_host(d_pad), keep)       # the padded input is left alone
    c2r.execute(A.PFX_BACKWARD, d_ref, scale=0.5)
    assert np.array_equal(to_host(d_out), to_host(d_ref)[:, :n])
    # argument checks
    c2c = P.Plan.c2c("f32", n, 4, 1, n)
    assert L.pfx_execute_r2c_oop(c2c.handle, d_real.data_ptr(), n, d_pad.data_ptr(), 1.0,
                                 None) == A.PFX_ERR_INVALID_ARG
    assert L.pfx_execute_r2c_oop(r2c.handle, d_real.data_ptr(), n - 1, d_pad.data_ptr(), 1.0,
                                 None) == A.PFX_ERR_INVALID_ARG
    assert L.pfx_execute_c2r_oop(r2c.handle, d_pad.data_ptr(), d_out.data_ptr(), n, 1.0,
                                 None) == A.PFX_ERR_INVALID_ARG
    odd = P.Plan.r2c("f32", 35, 4, 1, 36)              # no tiled kernel for an odd length
    assert L.pfx_execute_r2c_oop(odd.handle, d_real.data_ptr(), 36, d_pad.data_ptr(), 1.0,
                                 None) == A.PFX_ERR_UNSUPPORTED


def test_r2c_out_of_place_unsupported_touches_nothing():
    """An N-D plan whose real axis has no tiled kernel refuses the fused
    forms before running any stage, so the caller's copy + in-place fallback
    starts from untouched buffers."""
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    shape = (6, 8, 35)
    fft = P.ParaFaFT_R2C(list(shape), precision="f32", fast=True)
    spec = to_dev(O.uniform(fft.get_required_output_size(), seed=3).astype(np.float32))
    keep = to_host(spec).copy()
    real = to_dev(np.zeros(int(np.prod(shape)), np.float32))
    for direction in (A.PFX_FORWARD, A.PFX_BACKWARD):
        rc = A.load().pfx_ndplan_execute_r2c_oop(fft._nd.handle, direction, real.data_ptr(),
                                                 spec.data_ptr(), 1.0, None)
        assert rc == A.PFX_ERR_UNSUPPORTED
    assert np.array_equal(to_host(spec), keep)
    assert not to_host(real).any()
