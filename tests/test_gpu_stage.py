"""GPU parity of the batched strided 1D plans (execute_stage, R2C, C2R) against
the FP64 oracle, called through the C ABI.

Mirrors test/unit/test_hipfft_backend.hip (contiguous, strided, multi-stage,
same plan on another pointer, R2C/C2R in place), with the reference's
roundtrip checks replaced by direct comparison with the oracle's
FFTWBackend<double> restatement on the same seeded inputs.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import real_supported, rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

TILED = [2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 32, 40, 48, 60, 64, 80, 96, 100, 120, 128,
         160, 192, 240, 256, 320, 384, 512, 640, 768, 1000, 1024, 1280, 1536, 2048, 2560, 3072,
         4096]
FALLBACK = [1, 7, 11, 144, 97, 1001]
PRECS = ["f32", "f64"]


def _cdtype(prec):
    return np.complex64 if prec == "f32" else np.complex128


def _run_c2c(prec, n, batch, stride, dist, direction, buf, outer=1, outer_dist=0):
    import parafaft_amd as P
    plan = P.Plan.c2c(prec, n, batch, stride, dist, outer, outer_dist)
    d = to_dev(buf)
    plan.execute(direction, d)
    out = to_host(d)
    plan.destroy()
    return out


def _ref_c2c(buf, n, batch, stride, dist, sign, outer=1, outer_dist=0):
    ref = buf.astype(np.complex128).copy()
    for o in range(outer):
        view = ref[o * outer_dist:] if outer_dist else ref
        tmp = np.ascontiguousarray(view)
        O.c2c_batched(tmp, n, batch, stride, dist, sign)
        if outer_dist:
            ref[o * outer_dist:] = tmp
        else:
            ref = tmp
    return ref


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", TILED + FALLBACK)
def test_contiguous_rows(prec, n):
    batch = 37
    x = O.uniform_complex((batch * n,), seed=1000 + n).astype(_cdtype(prec))
    for direction, sign in ((0, -1), (1, +1)):
        y = _run_c2c(prec, n, batch, 1, n, direction, x)
        ref = _ref_c2c(x, n, batch, 1, n, sign)
        assert rel_l2(y, ref) <= tol(prec, n), (prec, n, direction, rel_l2(y, ref))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [2, 3, 8, 16, 48, 64, 80, 120, 256, 1024, 2048, 4096, 7, 100, 144,
                               384, 640, 1000, 1536, 2560, 3072])
@pytest.mark.parametrize("batch", [1, 5, 16, 67])
def test_batch_contiguous_columns(prec, n, batch):
    """Axis-0 layout: stride = batch, dist = 1 (parafaft_c2c.hpp:264-272)."""
    x = O.uniform_complex((n * batch,), seed=7 * n + batch).astype(_cdtype(prec))
    for direction, sign in ((0, -1), (1, +1)):
        y = _run_c2c(prec, n, batch, batch, 1, direction, x)
        ref = _ref_c2c(x, n, batch, batch, 1, sign)
        assert rel_l2(y, ref) <= tol(prec, n), (prec, n, batch, direction, rel_l2(y, ref))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(4, 64, 16), (3, 1024, 8), (5, 48, 33), (2, 7, 9)])
def test_middle_axis_single_launch(prec, shape):
    """Middle axis as one two-level launch == the reference's loop of
    leading_size execute_stage calls (parafaft_c2c.hpp:875-896)."""
    lead, n, trail = shape
    x = O.uniform_complex((lead * n * trail,), seed=lead + n + trail).astype(_cdtype(prec))
    y = _run_c2c(prec, n, trail, trail, 1, 0, x, outer=lead, outer_dist=n * trail)
    ref = x.astype(np.complex128).copy().reshape(lead, n * trail)
    for i in range(lead):
        row = np.ascontiguousarray(ref[i])
        O.c2c_batched(row, n, trail, trail, 1, -1)
        ref[i] = row
    assert rel_l2(y, ref) <= tol(prec, n)


@pytest.mark.parametrize("prec", PRECS)
def test_general_strides(prec):
    """Neither stride nor dist is 1; transforms interleaved with gaps."""
    n, batch, stride, dist = 64, 6, 7, 3
    size = (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size + 5,), seed=99).astype(_cdtype(prec))
    y = _run_c2c(prec, n, batch, stride, dist, 0, x)
    ref = _ref_c2c(x, n, batch, stride, dist, -1)
    assert rel_l2(y, ref) <= tol(prec, n)
    # untouched gaps stay bit-identical
    touched = np.zeros(x.size, bool)
    for b in range(batch):
        touched[b * dist + np.arange(n) * stride] = True
    assert np.array_equal(y[~touched], x[~touched])


@pytest.mark.parametrize("prec", PRECS)
def test_same_plan_other_pointer(prec):
    """New-array execute semantics (backend/fft_backend.hpp:107-111): a plan
    applies to any element-aligned pointer with its layout."""
    import parafaft_amd as P
    n, batch = 256, 8
    x = O.uniform_complex((3 * batch * n + 1,), seed=5).astype(_cdtype(prec))
    d = to_dev(x)
    plan = P.Plan.c2c(prec, n, batch, 1, n)
    esz = x.itemsize
    plan.execute(0, d.data_ptr() + esz * (batch * n + 1))
    y = to_host(d)
    ref = x.astype(np.complex128).copy()
    seg = np.ascontiguousarray(ref[batch * n + 1: 2 * batch * n + 1])
    O.c2c_batched(seg, n, batch, 1, n, -1)
    ref[batch * n + 1: 2 * batch * n + 1] = seg
    assert rel_l2(y, ref) <= tol(prec, n)
    assert np.array_equal(y[:batch * n + 1], x[:batch * n + 1])


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [4, 6, 8, 10, 16, 30, 32, 64, 80, 96, 120, 160, 256, 1024, 2048,
                               4096, 8192, 2, 3, 5, 7, 15, 99, 100, 200, 768, 2000, 3072, 6144])
def test_r2c_c2r_inplace(prec, n):
    """In-place padded R2C / C2R (fft_backend_fftw.hpp:439-549)."""
    import parafaft_amd as P
    batch = 9
    h = n // 2 + 1
    pd = 2 * h
    rdt = np.float32 if prec == "f32" else np.float64
    real = O.uniform(batch * n, seed=300 + n).reshape(batch, n).astype(rdt)
    pad = np.zeros((batch, pd), dtype=rdt)
    pad[:, :n] = real
    ref = pad.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, pd)
    refc = ref.view(np.complex128)
    if not real_supported(prec, n):
        with pytest.raises(P.PfxError):
            P.Plan.r2c(prec, n, batch, 1, pd)
        return
    plan = P.Plan.r2c(prec, n, batch, 1, pd)
    d = to_dev(pad)
    plan.execute(0, d)
    got = to_host(d)
    gotc = got.view(np.complex64 if prec == "f32" else np.complex128)
    assert rel_l2(gotc, refc) <= tol(prec, n), rel_l2(gotc, refc)
    # C2R of the GPU spectrum back to n * x
    c2r = P.Plan.c2r(prec, n, batch, 1, pd)
    c2r.execute(1, d)
    back = to_host(d)[:, :n]
    assert rel_l2(back, n * real.astype(np.float64)) <= 2 * tol(prec, n)
    # C2R against the oracle on the oracle's own spectrum
    ref2 = ref.copy()
    O.c2r_inplace(ref2, n, batch, 1, pd)
    d2 = to_dev(ref.astype(rdt))
    c2r.execute(1, d2)
    got2 = to_host(d2)[:, :n]
    assert rel_l2(got2, ref2[:, :n]) <= tol(prec, n)


@pytest.mark.parametrize("prec", PRECS)
def test_scaled_epilogue(prec):
    """Fused 1/N scaling == the user-side normalisation loop
    (example/c2c.cpp:27-31)."""
    import parafaft_amd as P
    n, batch = 512, 4
    x = O.uniform_complex((batch * n,), seed=11).astype(_cdtype(prec))
    d = to_dev(x)
    plan = P.Plan.c2c(prec, n, batch, 1, n)
    plan.execute(0, d)
    plan.execute(1, d, scale=1.0 / n)
    y = to_host(d)
    assert rel_l2(y, x) <= tol(prec, n)


@pytest.mark.parametrize("prec,tile", [("f32", 8), ("f32", 16), ("f32", 32), ("f32", 64),
                                       ("f64", 4), ("f64", 8), ("f64", 16)])
@pytest.mark.parametrize("n", [256, 512, 1024, 2048])
def test_column_tile_variants(prec, tile, n, monkeypatch):
    """Every batch-contiguous tile variant (plain LDS exchange, large tile,
    wide per-component exchange) against the oracle, with a ragged batch."""
    import parafaft_amd as P
    monkeypatch.setenv("PFX_COL_TILE", str(tile))
    batch = 3 * tile + 5
    x = O.uniform_complex((n * batch,), seed=n + tile).astype(_cdtype(prec))
    plan = P.Plan.c2c(prec, n, batch, batch, 1)
    name, per_block, _ = plan.describe()
    for direction, sign in ((0, -1), (1, +1)):
        d = to_dev(x)
        plan.execute(direction, d)
        y = to_host(d)
        ref = _ref_c2c(x, n, batch, batch, 1, sign)
        assert rel_l2(y, ref) <= tol(prec, n), (name, per_block, direction, rel_l2(y, ref))
    # middle-axis form with the same variant
    lead = 3
    xm = O.uniform_complex((lead * n * batch,), seed=7 + n).astype(_cdtype(prec))
    ym = _run_c2c(prec, n, batch, batch, 1, 0, xm, outer=lead, outer_dist=n * batch)
    ref = xm.astype(np.complex128).copy().reshape(lead, n * batch)
    for i in range(lead):
        row = np.ascontiguousarray(ref[i])
        O.c2c_batched(row, n, batch, batch, 1, -1)
        ref[i] = row
    assert rel_l2(ym, ref) <= tol(prec, n), name


@pytest.mark.parametrize("match,prec,n", [("wide:32, 32", "f32", 1024),
                                          ("wide:16, 16, 4", "f32", 1024),
                                          ("wide:16, 16, 8", "f32", 2048),
                                          ("wide:16, 16, 4", "f64", 1024),
                                          ("wide:8, 8, 8", "f64", 512)])
def test_named_column_kernels(match, prec, n, monkeypatch):
    """Kernels selected by radix description (the measurement override)."""
    import parafaft_amd as P
    monkeypatch.setenv("PFX_KERNEL_MATCH", match)
    batch = 70
    x = O.uniform_complex((n * batch,), seed=n + len(match)).astype(_cdtype(prec))
    plan = P.Plan.c2c(prec, n, batch, batch, 1)
    name = plan.describe()[0]
    assert match in name, name
    for direction, sign in ((0, -1), (1, +1)):
        d = to_dev(x)
        plan.execute(direction, d)
        ref = _ref_c2c(x, n, batch, batch, 1, sign)
        assert rel_l2(to_host(d), ref) <= tol(prec, n), (name, direction)


def test_wide_4096_column_kernel(monkeypatch):
    """The 4096-point 8-column wide variant (160 KiB LDS)."""
    import parafaft_amd as P
    monkeypatch.setenv("PFX_KERNEL_MATCH", "wide:16, 16, 16")
    n, batch = 4096, 37
    x = O.uniform_complex((n * batch,), seed=4096).astype(np.complex64)
    plan = P.Plan.c2c("f32", n, batch, batch, 1)
    assert "wide:16, 16, 16" in plan.describe()[0]
    for direction, sign in ((0, -1), (1, +1)):
        d = to_dev(x)
        plan.execute(direction, d)
        assert rel_l2(to_host(d), _ref_c2c(x, n, batch, batch, 1, sign)) <= tol("f32", n)


@pytest.mark.parametrize("kind", ["c2c", "r2c", "c2r"])
def test_empty_batch(kind):
    """A plan over zero transforms executes as a no-op (nothing launched,
    nothing touched)."""
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    x = np.arange(64, dtype=np.float32)
    d = to_dev(x)
    if kind == "c2c":
        plan = P.Plan.c2c("f32", 16, 0, 1, 16)
        plan.execute(A.PFX_FORWARD, d)
    elif kind == "r2c":
        plan = P.Plan.r2c("f32", 16, 0, 1, 18)
        plan.execute(A.PFX_FORWARD, d)
    else:
        plan = P.Plan.c2r("f32", 16, 0, 1, 18)
        plan.execute(A.PFX_BACKWARD, d)
    assert np.array_equal(to_host(d), x)
