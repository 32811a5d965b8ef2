"""Lengths beyond one kernel: the four-step C2C plan (n = n1 * n2, both
factors tiled), Bluestein's chirp convolution for lengths that split into
no tiled pair, and the long R2C / C2R plans built on them, against the
oracle.

hipFFT, which the backend replaces (backend/hipfft/fft_backend_hipfft.hpp:
294-314), plans any length; here a length with no tiled kernel beyond the
direct-DFT limit runs as: length-n2 transforms along j2 into a scratch
buffer, the twiddle W_n^(j1 k2), length-n1 transforms along j1, and a
transpose back (pfx_fourstep.hip).  Checked in rows, batch-contiguous
columns, general strides with gaps (every element the plan does not own
comes back bit-identical) and two-level batches, both directions, with a
fused scale; the FP32 tolerance is the tiled plans' 1e-6 log2 N."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import (chirp_length, long_split, rel_l2, tiled_lengths, to_dev, to_host,
                            tol)

pytestmark = pytest.mark.gpu

CASES = [  # (prec, n, layout); rows of 8192 and 16384 run one kernel
    ("f32", 32768, "row"), ("f32", 8192, "col"), ("f64", 32768, "row"),
    ("f32", 32768, "gen"), ("f64", 16384, "col"), ("f32", 6144, "row"),
    ("f64", 5120, "gen"), ("f32", 65536, "row"), ("f32", 12288, "col"),
    ("f64", 10240, "row"), ("f32", 1 << 20, "row"), ("f64", 24576, "row"),
]


def _geometry(n, layout):
    if layout == "row":
        return 3, 1, n + 5                       # batch, stride, dist (gaps)
    if layout == "col":
        return 5, 5, 1
    return 2, 3, 3 * (n - 1) + 1 + 7               # general strided rows with gaps


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_long_c2c_vs_oracle(case):
    import parafaft_amd as P
    prec, n, layout = case
    assert long_split(n) is not None
    batch, stride, dist = _geometry(n, layout)
    cdt = np.complex64 if prec == "f32" else np.complex128
    size = (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size,), seed=n + len(layout)).astype(cdt)
    owned = np.zeros(size, bool)
    idx = (np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride).ravel()
    owned[idx] = True
    plan = P.Plan.c2c(prec, n, batch, stride, dist)
    n1, n2 = long_split(n)
    assert plan.describe()[0].endswith(f"fourstep_n{n}_{n1}x{n2}"), plan.describe()
    for direction, sign, scale in ((0, -1, 1.0), (1, +1, 0.5)):
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, stride, dist, sign)
        ref *= scale
        d = to_dev(x)
        plan.execute(direction, d, scale=scale)
        got = to_host(d)
        e = rel_l2(got[owned], ref[owned])
        assert e <= tol(prec, n), (direction, e)
        assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8))
        del ref, got, d


def test_long_c2c_two_level_batch_and_roundtrip():
    """pfx_plan_create_c2c_many (outer slices run one after another on the
    plan's scratch) and backward(forward(x)) = n x."""
    import parafaft_amd as P
    n, inner, outer = 8192, 3, 2
    stride, dist, odist = 1, n, 3 * n + 11
    size = (outer - 1) * odist + (inner - 1) * dist + n
    x = O.uniform_complex((size,), seed=77).astype(np.complex64)
    plan = P.Plan.c2c("f32", n, inner, stride, dist, outer, odist)
    ref = x.astype(np.complex128).copy()
    for o in range(outer):
        sl = ref[o * odist:o * odist + (inner - 1) * dist + n]
        O.c2c_batched(sl, n, inner, stride, dist, -1)
        ref[o * odist:o * odist + sl.size] = sl
    d = to_dev(x)
    plan.execute(0, d)
    got = to_host(d)
    for o in range(outer):
        for b in range(inner):
            s = o * odist + b * dist
            assert rel_l2(got[s:s + n], ref[s:s + n]) <= tol("f32", n)
    plan.execute(1, d, scale=1.0 / n)
    back = to_host(d)
    for o in range(outer):
        for b in range(inner):
            s = o * odist + b * dist
            assert rel_l2(back[s:s + n], x[s:s + n]) <= 2 * tol("f32", n)


def test_long_c2c_measure_tuned():
    """Measure planning tunes the two sub-plans; the result is unchanged."""
    import parafaft_amd as P
    n, batch = 32768, 4
    x = O.uniform_complex((batch * n,), seed=5).astype(np.complex64)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, 1, n, -1)
    plan = P.Plan.c2c("f32", n, batch, 1, n)
    plan.tune(1, to_dev(np.zeros(batch * n, np.complex64)))
    d = to_dev(x)
    plan.execute(0, d)
    assert rel_l2(to_host(d), ref) <= tol("f32", n)


@pytest.mark.parametrize("prec,n", [("f32", 65536), ("f64", 16384), ("f32", 12288),
                                    ("f32", 1 << 18), ("f64", 8192 + 4096)])
def test_long_r2c_c2r_vs_oracle(prec, n):
    """Even real lengths whose half has no tiled kernel: the half-length
    complex FFT of the padded rows read as pairs, then the R2C post-pass;
    the C2R pre-pass, then the inverse."""
    import parafaft_amd as P
    batch = 3
    h = n // 2 + 1
    pd = 2 * h + 2
    rdt = np.float32 if prec == "f32" else np.float64
    real = O.uniform(batch * n, seed=n + 1).reshape(batch, n).astype(rdt)
    pad = np.zeros((batch, pd), dtype=rdt)
    pad[:, :n] = real
    ref = pad.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, pd)
    r2c = P.Plan.r2c(prec, n, batch, 1, pd)
    c2r = P.Plan.c2r(prec, n, batch, 1, pd)
    assert r2c.describe()[0].endswith(f"r2c_long_n{n}"), r2c.describe()
    assert c2r.describe()[0].endswith(f"c2r_long_n{n}"), c2r.describe()
    d = to_dev(pad)
    r2c.execute(0, d)
    got = to_host(d)
    cdt = np.complex64 if prec == "f32" else np.complex128
    refc = ref[:, :2 * h].copy().view(np.complex128)
    assert rel_l2(got[:, :2 * h].copy().view(cdt), refc) <= tol(prec, n)
    # C2R of the oracle's own spectrum against the oracle's C2R
    ref2 = ref.copy()
    O.c2r_inplace(ref2, n, batch, 1, pd)
    d2 = to_dev(ref.astype(rdt))
    c2r.execute(1, d2)
    assert rel_l2(to_host(d2)[:, :n], ref2[:, :n]) <= tol(prec, n)
    # round trip of the GPU's spectrum, scaled
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real.astype(np.float64)) <= 2 * tol(prec, n)


CHIRP_CASES = [  # lengths that split into no tiled pair: Bluestein
    ("f32", 8191, "row"), ("f32", 4099, "col"), ("f64", 2053, "gen"),
    ("f32", 10007, "gen"), ("f64", 7000, "row"), ("f32", 7000, "col"),
]


@pytest.mark.parametrize("case", CHIRP_CASES, ids=lambda c: "-".join(map(str, c)))
def test_chirp_c2c_vs_oracle(case):
    """Bluestein: the chirp premultiply into zero-padded rows of M >= 2n - 1,
    two M-point transforms around the filter's spectrum, the chirp
    postmultiply; rows, columns and strided layouts with gaps (elements the
    plan does not own come back bit-identical), both directions, a scale."""
    import parafaft_amd as P
    prec, n, layout = case
    assert long_split(n) is None
    batch, stride, dist = _geometry(n, layout)
    cdt = np.complex64 if prec == "f32" else np.complex128
    size = (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size,), seed=3 * n + len(layout)).astype(cdt)
    owned = np.zeros(size, bool)
    idx = (np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride).ravel()
    owned[idx] = True
    plan = P.Plan.c2c(prec, n, batch, stride, dist)
    M = chirp_length(prec, n)
    assert plan.describe()[0].endswith(f"bluestein_n{n}_m{M}"), plan.describe()
    for direction, sign, scale in ((0, -1, 1.0), (1, +1, 1.0 / n)):
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, stride, dist, sign)
        ref *= scale
        d = to_dev(x)
        plan.execute(direction, d, scale=scale)
        got = to_host(d)
        e = rel_l2(got[owned], ref[owned])
        assert e <= tol(prec, n), (direction, e)
        assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8))


@pytest.mark.parametrize("prec,n", [("f32", 4097), ("f64", 2049), ("f32", 8191),
                                    ("f32", 2 * 8191), ("f64", 6001)])
def test_odd_and_chirp_real_vs_oracle(prec, n):
    """Odd real lengths beyond the direct-DFT limit run the whole row as a
    C2C on scratch rows (R2C: the reals with zero imaginary parts; C2R: the
    Hermitian extension); an even length whose half has no tiled pair runs
    its half through the chirp convolution."""
    import parafaft_amd as P
    batch = 3
    h = n // 2 + 1
    pd = 2 * h + 2
    rdt = np.float32 if prec == "f32" else np.float64
    real = O.uniform(batch * n, seed=n + 7).reshape(batch, n).astype(rdt)
    pad = np.zeros((batch, pd), dtype=rdt)
    pad[:, :n] = real
    ref = pad.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, pd)
    r2c = P.Plan.r2c(prec, n, batch, 1, pd)
    c2r = P.Plan.c2r(prec, n, batch, 1, pd)
    d = to_dev(pad)
    r2c.execute(0, d)
    got = to_host(d)
    cdt = np.complex64 if prec == "f32" else np.complex128
    refc = ref[:, :2 * h].copy().view(np.complex128)
    assert rel_l2(got[:, :2 * h].copy().view(cdt), refc) <= tol(prec, n)
    ref2 = ref.copy()
    O.c2r_inplace(ref2, n, batch, 1, pd)
    d2 = to_dev(ref.astype(rdt))
    c2r.execute(1, d2)
    assert rel_l2(to_host(d2)[:, :n], ref2[:, :n]) <= tol(prec, n)
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real.astype(np.float64)) <= 2 * tol(prec, n)


@pytest.mark.parametrize("fast", [True, False])
def test_driver_with_a_long_axis(fast):
    """A 3 x 8192 C2C and a 2 x 16384 R2C through the one-rank drivers (the
    backend's create_stage_plan gets the long length: one-kernel rows, a
    four-step batch-contiguous axis of 3 and 2)."""
    import parafaft_amd as P
    shape = [3, 8192]
    N = int(np.prod(shape))
    x = O.uniform_complex(tuple(shape), seed=9).astype(np.complex64)
    ref = O.fftn(x.astype(np.complex128))
    fft = P.ParaFaFT_C2C(shape, precision="f32", fast=fast)
    d = to_dev(x)
    fft.forward(d)
    assert rel_l2(to_host(d), ref) <= tol("f32", N)
    fft.backward(d, scale=1.0 / N)
    assert rel_l2(to_host(d), x) <= 2 * tol("f32", N)
    shape = [2, 16384]
    N = int(np.prod(shape))
    xr = O.uniform(N, seed=10).reshape(shape).astype(np.float32)
    refr = O.rfftn_padded(xr.astype(np.float64))
    fr = P.ParaFaFT_R2C(shape, precision="f32", fast=fast)
    h = shape[-1] // 2 + 1
    pad = np.zeros(fr.get_required_output_size(), np.float32)
    pad[:2 * 2 * h].reshape(2, 2 * h)[:, :shape[-1]] = xr
    dr = to_dev(pad)
    fr.forward_in_place(dr)
    got = to_host(dr)[:2 * 2 * h].copy().view(np.complex64).reshape(2, h)
    assert rel_l2(got, refr) <= tol("f32", N)


def test_long_refusals():
    """Lengths beyond the chirp limit that do not split are refused; a long
    plan has no segmented or out-of-place real form and no persistent body;
    in a batch it runs on its own (direct) beside the grouped launches."""
    import ctypes as C
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    from parafaft_amd.batch import Batch
    big = (1 << 23) + 1   # splits into no tiled pair and exceeds the chirp limit
    with pytest.raises(P.PfxError) as e:
        P.Plan.c2c("f32", big, 2, 1, big)
    assert e.value.status == 2
    with pytest.raises(P.PfxError):
        P.Plan.r2c("f32", 2 * big, 2, 1, 2 * (big + 1))
    n = 32768  # a four-step plan in rows (8192 and 16384 rows run one kernel)
    plan = P.Plan.c2c("f32", n, 2, 1, n)
    assert "fourstep" in plan.describe()[0]
    d = to_dev(np.zeros(2 * n, np.complex64))
    seg = (C.c_void_p * 1)(d.data_ptr())
    st = (C.c_int * 1)(0)
    cnt = (C.c_int * 1)(n)
    m = A.SegMap(1, seg, st, cnt, 1)
    rc = A.load().pfx_execute_segmented(plan.handle, 0, d.data_ptr(), C.byref(m), None, 1.0, None)
    assert rc == A.PFX_ERR_UNSUPPORTED
    # a batch with one long transform: grouped schedule matches the oracle
    x0 = O.uniform_complex((n,), seed=1).astype(np.complex64)
    x1 = O.uniform_complex((64, 64), seed=2).astype(np.complex64)
    buf = to_dev(np.concatenate([x0, x1.ravel()]))
    b = Batch("f32", [("c2c", (n,), buf.data_ptr()), ("c2c", (64, 64), buf.data_ptr() + n * 8)])
    b.execute(0)
    got = to_host(buf)
    assert rel_l2(got[:n], O.fftn(x0.astype(np.complex128))) <= tol("f32", n)
    assert rel_l2(got[n:].reshape(64, 64), O.fftn(x1.astype(np.complex128))) <= tol("f32", 4096)
    with pytest.raises(P.PfxError):
        b.execute_persistent(0)


@pytest.mark.parametrize("prec,n", [("f32", 7), ("f32", 144), ("f32", 4093), ("f64", 2047),
                                    ("f64", 99)])
def test_direct_dft_kept_under_the_knob(prec, n, monkeypatch):
    """PFX_LONG_MIN (a measurement knob) keeps the direct O(n^2) DFT below
    its value, as the allocation-free entry point always does: C2C and
    R2C / C2R against the oracle, so the fallback kernel stays covered."""
    import parafaft_amd as P
    monkeypatch.setenv("PFX_LONG_MIN", str(1 << 30))
    batch = 3
    cdt = np.complex64 if prec == "f32" else np.complex128
    x = O.uniform_complex((batch * n,), seed=n + 11).astype(cdt)
    plan = P.Plan.c2c(prec, n, batch, 1, n)
    assert "direct_dft" in plan.describe()[0], plan.describe()
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, 1, n, -1)
    d = to_dev(x)
    plan.execute(0, d)
    assert rel_l2(to_host(d), ref) <= tol(prec, n)
    h = n // 2 + 1
    pd = 2 * h
    rdt = np.float32 if prec == "f32" else np.float64
    real = O.uniform(batch * n, seed=n + 12).reshape(batch, n).astype(rdt)
    pad = np.zeros((batch, pd), rdt)
    pad[:, :n] = real
    refr = pad.astype(np.float64).copy()
    O.r2c_inplace(refr, n, batch, 1, pd)
    r2c = P.Plan.r2c(prec, n, batch, 1, pd)
    if n % 2 == 0 and n // 2 in tiled_lengths():
        return   # a tiled R2C kernel exists
    assert "direct_dft" in r2c.describe()[0], r2c.describe()
    d = to_dev(pad)
    r2c.execute(0, d)
    got = to_host(d).copy().view(cdt)
    assert rel_l2(got, refr.view(np.complex128)) <= tol(prec, n)
    c2r = P.Plan.c2r(prec, n, batch, 1, pd)
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real.astype(np.float64)) <= 2 * tol(prec, n)


@pytest.mark.parametrize("prec,n", [("f32", 7), ("f32", 144), ("f64", 99), ("f32", 1001)])
def test_short_lengths_without_a_kernel_run_long_plans(prec, n):
    """By default a length with no tiled kernel runs a long plan (four-step
    when it splits into two tiled lengths, else the chirp convolution), not
    the direct DFT."""
    import parafaft_amd as P
    plan = P.Plan.c2c(prec, n, 2, 1, n)
    name = plan.describe()[0]
    assert ("fourstep" in name) == (long_split(n) is not None), name
    assert "direct_dft" not in name


def _random_lengths(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(5, 12000))
        prec = "f64" if rng.random() < 0.4 else "f32"
        layout = ("row", "col", "gen")[int(rng.integers(0, 3))]
        out.append((i, prec, n, layout))
    return out


@pytest.mark.parametrize("case", _random_lengths(20261204, 24),
                         ids=lambda c: "-".join(map(str, c)))
def test_random_lengths_vs_oracle(case):
    """Seeded random lengths 5..12000 (tiled, four-step or chirp, whichever
    the length takes), random layout and precision: forward against the
    oracle, elements the plan does not own untouched."""
    import parafaft_amd as P
    i, prec, n, layout = case
    batch, stride, dist = _geometry(n, layout)
    cdt = np.complex64 if prec == "f32" else np.complex128
    size = (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size,), seed=9000 + i).astype(cdt)
    owned = np.zeros(size, bool)
    idx = (np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride).ravel()
    owned[idx] = True
    plan = P.Plan.c2c(prec, n, batch, stride, dist)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, stride, dist, -1)
    d = to_dev(x)
    plan.execute(0, d)
    got = to_host(d)
    assert rel_l2(got[owned], ref[owned]) <= tol(prec, n), (plan.describe(), rel_l2(got[owned],
                                                                                  ref[owned]))
    assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8))


def test_chirp_and_odd_real_batches_run_in_slices():
    """Batches whose chirp (odd-real) scratch rows exceed 256 MB run in
    slices of rows through one bounded scratch: 20000 transforms of 1001
    points in columns (M = 2048: two slices, the second short) and 20000
    odd real rows of 1001 (FP64), against the oracle."""
    import parafaft_amd as P
    n, batch = 1001, 20000
    x = O.uniform_complex((n * batch,), seed=31).astype(np.complex64)
    plan = P.Plan.c2c("f32", n, batch, batch, 1)
    assert "bluestein" in plan.describe()[0]
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, batch, 1, -1)
    d = to_dev(x)
    plan.execute(0, d)
    assert rel_l2(to_host(d), ref) <= tol("f32", n)
    del ref, d
    n, batch = 1001, 20000
    h = n // 2 + 1
    pd = 2 * h
    real = O.uniform(batch * n, seed=32).reshape(batch, n)
    pad = np.zeros((batch, pd))
    pad[:, :n] = real
    refr = pad.copy()
    O.r2c_inplace(refr, n, batch, 1, pd)
    r2c = P.Plan.r2c("f64", n, batch, 1, pd)
    c2r = P.Plan.c2r("f64", n, batch, 1, pd)
    d = to_dev(pad)
    r2c.execute(0, d)
    assert rel_l2(to_host(d).copy().view(np.complex128), refr.view(np.complex128)) <= tol("f64", n)
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real) <= 2 * tol("f64", n)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("n", [8192, 1 << 20])
def test_long_every_sub_plan_kernel(prec, layout, n, tmp_path, monkeypatch):
    """Every kernel candidate of both tiled sub-plans of a four-step plan
    (8192 = 64 x 128 in columns, 32768 = 128 x 256 in rows, 2^20 = 1024 x
    1024), forced in turn through a tuning-cache line: the first pass runs
    out of place into the plan's scratch, so every kernel must load from the
    caller's array; both directions against the oracle, gaps untouched.
    Rows take the four-pass form here (their default two-pass form has no
    sub-plans)."""
    from tests.gpu_util import forced_sub_plan_kernels
    if layout == "row":
        monkeypatch.setenv("PFX_FOURSTEP_PASSES", "4")
    if n == 8192 and layout == "row":
        n = 32768  # rows of 8192 run one kernel
    batch, stride, dist = _geometry(n, layout)
    if n > 8192:
        batch = 2
        stride, dist = (1, n + 5) if layout == "row" else (2, 1)
    size = (batch - 1) * dist + (n - 1) * stride + 1
    cdt = np.complex128 if prec == "f64" else np.complex64
    x = O.uniform_complex((size,), seed=77).astype(cdt)
    refs = {}
    for direction, sign in ((0, -1), (1, +1)):
        r = x.astype(np.complex128).copy()
        O.c2c_batched(r, n, batch, stride, dist, sign)
        refs[direction] = r
    touched = refs[0] != x
    forced = 0
    for key, name, plan in forced_sub_plan_kernels(prec, n, batch, stride, dist,
                                                   str(tmp_path / "cache.txt")):
        for direction in (0, 1):
            d = to_dev(x)
            plan.execute(direction, d)
            got = to_host(d)
            assert np.all(got[~touched] == x[~touched]), (key, name)
            err = rel_l2(got, refs[direction])
            assert err <= tol(prec, n), (prec, layout, key, name, direction, err)
        forced += 1
    assert forced >= 1, forced  # at least the kernel of each tiled sub-plan
