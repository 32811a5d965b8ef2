"""Chunked N-D schedule (pfx_ndplan_set_chunk / pfx_ndplan_chunk): every
stage but axis 0 runs on a chunk of axis-0 planes at a time.  Same kernels as
the stage-by-stage schedule, so results must be bit-identical; forward is
also checked against the oracle and backward(scale 1/N) against the input."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, torch

pytestmark = pytest.mark.gpu

CASES = [("c2c", (32, 24, 40), 8), ("c2c", (16, 16, 16, 8), 4), ("r2c", (24, 20, 30), 6),
         ("r2c", (8, 6, 10, 16), 2), ("c2c", (12, 64, 64), 3)]


def _nd(prec, kind, shape):
    from parafaft_amd import _capi as A
    h = C.c_void_p()
    A.check(A.load().pfx_ndplan_create(C.byref(h), prec, A.PFX_C2C if kind == "c2c" else A.PFX_R2C,
                                       len(shape), (C.c_int * len(shape))(*shape)), "create")
    return h.value


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("prec_name", ["f32", "f64"])
def test_chunked_matches_stagewise(case, prec_name):
    from parafaft_amd import _capi as A
    L = A.load()
    kind, shape, P = CASES[case]
    prec = A.PFX_F32 if prec_name == "f32" else A.PFX_F64
    cdt = np.complex64 if prec == A.PFX_F32 else np.complex128
    rdt = np.float32 if prec == A.PFX_F32 else np.float64
    if kind == "c2c":
        x = O.uniform_complex(shape, seed=300 + case).astype(cdt)
        host = x
        ref = O.fftn(x.astype(np.complex128))
    else:
        x = O.uniform(int(np.prod(shape)), seed=300 + case).reshape(shape).astype(rdt)
        n = shape[-1]
        h = n // 2 + 1
        host = np.zeros(shape[:-1] + (2 * h,), rdt)
        host[..., :n] = x
        ref = O.rfftn_padded(x.astype(np.float64))
    N = int(np.prod(shape))
    outs = {}
    for planes in (0, P):
        nd = _nd(prec, kind, shape)
        A.check(L.pfx_ndplan_set_chunk(nd, planes), "set_chunk")
        got = C.c_int()
        A.check(L.pfx_ndplan_chunk(nd, C.byref(got)), "chunk")
        assert got.value == planes
        d = to_dev(host)
        A.check(L.pfx_ndplan_execute(nd, A.PFX_FORWARD, d.data_ptr(), 1.0, None), "fwd")
        fwd = to_host(d).copy()
        A.check(L.pfx_ndplan_execute(nd, A.PFX_BACKWARD, d.data_ptr(), 1.0 / N, None), "bwd")
        outs[planes] = (fwd, to_host(d).copy())
        L.pfx_ndplan_destroy(nd)
    assert np.array_equal(outs[0][0].view(np.uint8), outs[P][0].view(np.uint8))
    assert np.array_equal(outs[0][1].view(np.uint8), outs[P][1].view(np.uint8))
    fwd, back = outs[P]
    tol = 1e-13 if prec == A.PFX_F64 else 1e-6 * np.log2(N)
    if kind == "c2c":
        assert rel_l2(fwd, ref) <= tol
        assert rel_l2(back, x) <= tol
    else:
        cplx = fwd.view(cdt).reshape(ref.shape)
        assert rel_l2(cplx, ref) <= tol
        assert rel_l2(back[..., :shape[-1]], x) <= tol


def test_chunk_rules():
    """Default: on only for ndim >= 3 arrays beyond twice the budget, with
    the largest divisor of n0 whose planes fit 256 MB; set_chunk validates."""
    from parafaft_amd import _capi as A
    L = A.load()
    got = C.c_int()
    for shape, prec, want in [((512, 512, 512), A.PFX_F32, 128), ((512, 512, 512), A.PFX_F64, 64),
                              ((1024, 1024, 1024), A.PFX_F32, 32), ((256, 256, 256), A.PFX_F32, 0),
                              ((4096, 4096), A.PFX_F32, 0)]:
        nd = _nd(prec, "c2c", shape)
        A.check(L.pfx_ndplan_chunk(nd, C.byref(got)), "chunk")
        assert got.value == want, (shape, prec, got.value)
        L.pfx_ndplan_destroy(nd)
    nd = _nd(A.PFX_F32, "c2c", (12, 8, 8))
    assert L.pfx_ndplan_set_chunk(nd, 5) == A.PFX_ERR_INVALID_ARG
    assert L.pfx_ndplan_set_chunk(nd, 12) == A.PFX_ERR_INVALID_ARG
    assert L.pfx_ndplan_set_chunk(nd, 4) == A.PFX_SUCCESS
    L.pfx_ndplan_destroy(nd)
