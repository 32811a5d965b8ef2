"""pfx_execute_c2c_noalloc: one C2C transform whose twiddle tables live in the
library's fixed device scratch array (no allocation).  Two calls in a row,
the second smaller, each bit-identical to the planned transform
(pfx_plan_create_c2c + pfx_execute_scaled, default kernel) and within the
FP32 / FP64 bound of the oracle; lengths past the scratch maximum are
refused."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

CASES = [  # (precision, length, batch, stride, dist, direction, scale): the second is smaller
    ("f32", 1024, 37, 37, 1, 0, 1.0),
    ("f64", 48, 5, 1, 50, 1, 1.0 / 48),
]


def test_noalloc_matches_planned_transform_twice():
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    from parafaft_amd import backend as B
    L = A.load()
    B.tune_cache_clear()   # the planned reference takes its default kernel too
    for prec, n, batch, stride, dist, direction, scale in CASES:
        cdt = np.complex64 if prec == "f32" else np.complex128
        size = (batch - 1) * dist + (n - 1) * stride + 1
        x = O.uniform_complex((size,), seed=n + batch).astype(cdt)
        d_plan, d_free = to_dev(x), to_dev(x)
        plan = P.Plan.c2c(prec, n, batch, stride, dist)
        plan.execute(direction, d_plan, scale=scale)
        rc = L.pfx_execute_c2c_noalloc(A.PFX_F32 if prec == "f32" else A.PFX_F64, n, batch,
                                       stride, dist, direction, C.c_void_p(d_free.data_ptr()),
                                       scale, None)
        A.check(rc, "pfx_execute_c2c_noalloc")
        got, want = to_host(d_free), to_host(d_plan)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (prec, n)
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, stride, dist, -1 if direction == 0 else +1)
        ref *= scale
        idx = (np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride).ravel()
        assert rel_l2(got[idx], ref[idx]) <= tol(prec, n), (prec, n)


def test_noalloc_direct_dft_for_lengths_without_a_kernel():
    """A length with no tiled kernel runs the direct DFT here (the planned
    path takes a long plan, which needs scratch): against the oracle."""
    from parafaft_amd import _capi as A
    L = A.load()
    for prec, n, batch in (("f32", 144, 3), ("f64", 7, 2)):
        cdt = np.complex64 if prec == "f32" else np.complex128
        x = O.uniform_complex((batch * n,), seed=n).astype(cdt)
        d = to_dev(x)
        rc = L.pfx_execute_c2c_noalloc(A.PFX_F32 if prec == "f32" else A.PFX_F64, n, batch, 1, n,
                                       0, C.c_void_p(d.data_ptr()), 1.0, None)
        A.check(rc, "pfx_execute_c2c_noalloc")
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, 1, n, -1)
        assert rel_l2(to_host(d), ref) <= tol(prec, n), (prec, n)


def test_noalloc_refuses_lengths_past_the_scratch():
    from parafaft_amd import _capi as A
    L = A.load()
    d = to_dev(np.zeros(8192, np.complex64))
    rc = L.pfx_execute_c2c_noalloc(A.PFX_F32, 8192, 1, 1, 8192, 0, C.c_void_p(d.data_ptr()),
                                   1.0, None)
    assert rc == A.PFX_ERR_INVALID_ARG
    assert b"scratch maximum" in L.pfx_last_error_message()
