"""Split axis 0 of the one-GPU N-D plan (pfx_ndplan_set_split): n0 = q G,
the G-point DFTs along axis 0 fused with the last-axis pass into a scratch,
the middle axes there, then the q-point DFTs back into the array.  Forward
and backward are checked against the oracle's N-D transform (sign -1 and +1)
for every built (precision, last-axis length, G), in 3-D and 4-D; the
factors that have no fused kernel, and plans the schedule does not apply to,
must be refused.  The tuned and the forced schedule match the oracle, HIP
graph replays are bit-identical to eager runs, and at 1024^3 the split
schedule agrees with the chunked one and round-trips."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, tol, to_dev, to_host, torch

pytestmark = pytest.mark.gpu

# (shape, chunk planes, G values)
CASES = [((32, 8, 1024), 8, (4, 8)), ((32, 4, 1024), 16, (4, 8, 16)),
         ((64, 16, 512), 16, (4, 8)), ((16, 4, 2048), 8, (4, 8)),
         ((16, 2, 4, 1024), 8, (4, 8)), ((24, 8, 1024), 12, (4,))]
# last-axis length -> G values with a fused kernel
BUILT = {"f32": {512: (4, 8), 1024: (4, 8, 16), 2048: (4, 8)},
         "f64": {512: (4, 8), 1024: (4, 8)}}


def _lib():
    from parafaft_amd import _capi as A
    return A, A.load()


def _nd(prec, shape, kind=None):
    A, L = _lib()
    h = C.c_void_p()
    A.check(L.pfx_ndplan_create(C.byref(h), prec, A.PFX_C2C if kind is None else kind, len(shape),
                                (C.c_int * len(shape))(*shape)), "pfx_ndplan_create")
    return h.value


def _exec(h, direction, d, scale=1.0, stream=None):
    A, L = _lib()
    A.check(L.pfx_ndplan_execute(h, direction, C.c_void_p(d.data_ptr()), scale, stream),
            "pfx_ndplan_execute")


def _split(h):
    A, L = _lib()
    g = C.c_int()
    A.check(L.pfx_ndplan_split(h, C.byref(g)), "pfx_ndplan_split")
    return g.value


PARAMS = [(c, p, g) for c, (shape, _, gs) in enumerate(CASES) for p in ("f32", "f64")
          for g in gs]


@pytest.mark.parametrize("case,prec_name,g", PARAMS)
def test_split_matches_oracle(case, prec_name, g):
    A, L = _lib()
    shape, planes, _ = CASES[case]
    prec = A.PFX_F32 if prec_name == "f32" else A.PFX_F64
    cdt = np.complex64 if prec_name == "f32" else np.complex128
    h = _nd(prec, shape)
    try:
        A.check(L.pfx_ndplan_set_chunk(h, planes), "pfx_ndplan_set_chunk")
        rc = L.pfx_ndplan_set_split(h, g)
        if g not in BUILT[prec_name].get(shape[-1], ()):
            assert rc == A.PFX_ERR_UNSUPPORTED, "a factor without a fused kernel must be refused"
            return
        A.check(rc, "pfx_ndplan_set_split")
        assert _split(h) == g
        x = O.uniform_complex(shape, seed=700 + case).astype(cdt)
        N = x.size
        d = to_dev(x)
        _exec(h, A.PFX_FORWARD, d)
        y = to_host(d)
        assert rel_l2(y, O.fftn(x.astype(np.complex128))) <= tol(prec_name, N)
        # backward of the spectrum with the 1/N scale fused returns the input
        _exec(h, A.PFX_BACKWARD, d, 1.0 / N)
        assert rel_l2(to_host(d), x) <= tol(prec_name, N)
        # backward on its own (sign +1, unscaled) against the oracle
        d = to_dev(x)
        _exec(h, A.PFX_BACKWARD, d)
        assert rel_l2(to_host(d), O.fftn(x.astype(np.complex128), sign=+1)) <= tol(prec_name, N)
    finally:
        L.pfx_ndplan_destroy(h)


def test_split_refusals():
    A, L = _lib()
    # stage-by-stage plan (no chunk): refused
    h = _nd(A.PFX_F32, (32, 8, 1024))
    try:
        assert L.pfx_ndplan_set_split(h, 8) == A.PFX_ERR_UNSUPPORTED
        A.check(L.pfx_ndplan_set_chunk(h, 8), "pfx_ndplan_set_chunk")
        # G must divide the chunk and n0
        assert L.pfx_ndplan_set_split(h, 16) == A.PFX_ERR_INVALID_ARG
        assert L.pfx_ndplan_set_split(h, 3) == A.PFX_ERR_INVALID_ARG
        assert L.pfx_ndplan_set_split(h, -1) == A.PFX_ERR_INVALID_ARG
        A.check(L.pfx_ndplan_set_split(h, 8), "pfx_ndplan_set_split")
        sp = C.c_void_p()
        A.check(L.pfx_ndplan_split_stage(h, C.byref(sp)), "pfx_ndplan_split_stage")
        assert sp.value
        # a new chunk size turns the split schedule off
        A.check(L.pfx_ndplan_set_chunk(h, 16), "pfx_ndplan_set_chunk")
        assert _split(h) == 0
        assert L.pfx_ndplan_split_stage(h, C.byref(sp)) == A.PFX_ERR_INVALID_ARG
        A.check(L.pfx_ndplan_set_split(h, 8), "pfx_ndplan_set_split")
        A.check(L.pfx_ndplan_set_split(h, 0), "pfx_ndplan_set_split")
        assert _split(h) == 0
    finally:
        L.pfx_ndplan_destroy(h)
    # R2C plans, split-precision forms and 2-D plans: refused
    for prec, kind, shape in ((A.PFX_F32, A.PFX_R2C, (32, 8, 1024)),
                              (A.PFX_F32_SPLIT_F16, A.PFX_C2C, (32, 8, 1024)),
                              (A.PFX_F32, A.PFX_C2C, (64, 1024))):
        h = _nd(prec, shape, kind)
        try:
            if len(shape) >= 3:
                A.check(L.pfx_ndplan_set_chunk(h, 8), "pfx_ndplan_set_chunk")
            assert L.pfx_ndplan_set_split(h, 8) == A.PFX_ERR_UNSUPPORTED
        finally:
            L.pfx_ndplan_destroy(h)


@pytest.mark.parametrize("schedule", ["tuned", "split"])
def test_split_tuned_plan_matches_oracle(schedule, monkeypatch):
    A, L = _lib()
    shape = (32, 8, 1024)
    if schedule == "split":
        monkeypatch.setenv("PFX_ND_SCHEDULE", "split")
    h = _nd(A.PFX_F32, shape)
    try:
        A.check(L.pfx_ndplan_set_chunk(h, 16), "pfx_ndplan_set_chunk")
        scratch = to_dev(np.zeros(shape, np.complex64))
        A.check(L.pfx_ndplan_tune(h, A.PFX_PLAN_MEASURE, C.c_void_p(scratch.data_ptr()), None),
                "pfx_ndplan_tune")
        torch().cuda.synchronize()
        g = _split(h)
        if schedule == "split":
            assert g in (4, 8, 16)
        x = O.uniform_complex(shape, seed=77).astype(np.complex64)
        d = to_dev(x)
        _exec(h, A.PFX_FORWARD, d)
        assert rel_l2(to_host(d), O.fftn(x.astype(np.complex128))) <= tol("f32", x.size)
    finally:
        L.pfx_ndplan_destroy(h)


def test_split_graph_replays_bit_identical():
    A, L = _lib()
    t = torch()
    shape = (32, 8, 1024)
    h = _nd(A.PFX_F32, shape)
    try:
        A.check(L.pfx_ndplan_set_chunk(h, 16), "pfx_ndplan_set_chunk")
        A.check(L.pfx_ndplan_set_split(h, 8), "pfx_ndplan_set_split")
        x = O.uniform_complex(shape, seed=78).astype(np.complex64)
        eager = to_dev(x)
        _exec(h, A.PFX_FORWARD, eager)
        _exec(h, A.PFX_BACKWARD, eager, 0.5)
        ref = to_host(eager)
        s = t.cuda.Stream()
        d = to_dev(x)
        g = C.c_void_p()
        t.cuda.synchronize()
        A.check(L.pfx_graph_begin_capture(C.c_void_p(s.cuda_stream)), "begin capture")
        try:
            _exec(h, A.PFX_FORWARD, d, 1.0, C.c_void_p(s.cuda_stream))
            _exec(h, A.PFX_BACKWARD, d, 0.5, C.c_void_p(s.cuda_stream))
        finally:
            A.check(L.pfx_graph_end_capture(C.c_void_p(s.cuda_stream), C.byref(g)), "end capture")
        try:
            for _ in range(3):
                d.copy_(to_dev(x))
                t.cuda.synchronize()
                A.check(L.pfx_graph_launch(g, C.c_void_p(s.cuda_stream)), "pfx_graph_launch")
                s.synchronize()
                assert np.array_equal(to_host(d), ref)
                # eager work between replays
                e2 = to_dev(x)
                _exec(h, A.PFX_FORWARD, e2)
                t.cuda.synchronize()
        finally:
            L.pfx_graph_destroy(g)
    finally:
        L.pfx_ndplan_destroy(h)


def test_split_1024cube_agrees_with_chunked():
    """1024^3 complex<float>: the split schedule (G = 8) against the chunked
    one at rel-L2 <= 2e-6, and its backward with the fused 1/N^3 scale
    returns the input at rel-L2 <= 2e-6."""
    A, L = _lib()
    t = torch()
    n = 1024
    g = t.Generator(device="cuda")
    g.manual_seed(31)
    x = t.view_as_complex(t.rand((n, n, n, 2), generator=g, device="cuda") * 2 - 1)
    h = _nd(A.PFX_F32, (n, n, n))
    try:
        chunk = C.c_int()
        A.check(L.pfx_ndplan_chunk(h, C.byref(chunk)), "pfx_ndplan_chunk")
        assert chunk.value % 8 == 0
        y = x.clone()
        _exec(h, A.PFX_FORWARD, y)
        A.check(L.pfx_ndplan_set_split(h, 8), "pfx_ndplan_set_split")
        z = x.clone()
        _exec(h, A.PFX_FORWARD, z)
        t.cuda.synchronize()
        num = den = 0.0
        for i in range(0, n, 64):
            r = y[i:i + 64].to(t.complex128)
            num += float(t.linalg.vector_norm(z[i:i + 64].to(t.complex128) - r) ** 2)
            den += float(t.linalg.vector_norm(r) ** 2)
        assert (num / den) ** 0.5 <= 2e-6
        del y
        _exec(h, A.PFX_BACKWARD, z, 1.0 / n ** 3)
        t.cuda.synchronize()
        num = den = 0.0
        for i in range(0, n, 64):
            r = x[i:i + 64].to(t.complex128)
            num += float(t.linalg.vector_norm(z[i:i + 64].to(t.complex128) - r) ** 2)
            den += float(t.linalg.vector_norm(r) ** 2)
        assert (num / den) ** 0.5 <= 2e-6
    finally:
        L.pfx_ndplan_destroy(h)
