"""The two-pass four-step of rows (pfx_fourstep.hip: fourstep_rows_pass1 /
_pass2): the length-n2 DFTs with the twiddle W_n^(j1 k2) fused, into the
plan's scratch at the same offsets, then the length-n1 DFTs stored
transposed into the caller's array with the caller's scale.  Against the
oracle for the factor pairs the fused kernels cover (powers of two 64..2048
and the tiled smooth lengths 80..3072),
on contiguous, gapped, strided and two-level rows, both directions;
elements the plan does not own come back bit-identical; the four-pass form
(PFX_FOURSTEP_PASSES=4) gives the oracle's result on the same inputs."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import long_split, rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

CASES = [  # (prec, n): n1 x n2 by the most even split
    ("f32", 1 << 15), ("f64", 1 << 15),    # 128 x 256
    ("f32", 1 << 16), ("f64", 1 << 17),    # 256 x 256, 256 x 512
    ("f32", 1 << 18), ("f64", 1 << 19),    # 512 x 512, 512 x 1024
    ("f32", 1 << 20), ("f64", 1 << 20),    # 1024 x 1024
    ("f32", 1 << 21), ("f64", 1 << 21),    # 1024 x 2048
    ("f32", 1 << 22),                      # 2048 x 2048
    ("f32", 6144), ("f64", 12288),         # 64 x 96, 96 x 128
    ("f32", 10240), ("f64", 24576),        # 80 x 128, 128 x 192
    ("f32", 15 * 4096), ("f32", 3 << 16),  # 240 x 256, 384 x 512
    ("f64", 1536 * 2048), ("f32", 10 ** 6),  # 1536 x 2048, 1000 x 1000
    ("f64", 10 ** 5),                      # 100 x 1000
]
GEOMS = {  # batch, stride, dist(n), outer, outer gap
    "contiguous": (3, 1, lambda n: n, 1, 0),
    "gapped": (2, 1, lambda n: n + 5, 2, 17),
    "strided": (2, 3, lambda n: 3 * (n - 1) + 1 + 7, 1, 0),
}


def _geom(n, name):
    batch, stride, dist, outer, gap = GEOMS[name]
    dist = dist(n)
    span = (batch - 1) * dist + (n - 1) * stride + 1
    return batch, stride, dist, outer, span + gap, (outer - 1) * (span + gap) + span


def _oracle(x, n, batch, stride, dist, outer, odist, sign):
    r = x.astype(np.complex128).copy()
    for o in range(outer):
        v = r[o * odist:].copy()
        O.c2c_batched(v, n, batch, stride, dist, sign)
        r[o * odist:] = v
    return r


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
@pytest.mark.parametrize("geom", list(GEOMS))
def test_rows_two_pass_vs_oracle(case, geom, monkeypatch):
    import parafaft_amd as P
    prec, n = case
    if n >= (1 << 21) and geom != "contiguous":
        pytest.skip("one layout at the largest lengths keeps the oracle's time bounded")
    batch, stride, dist, outer, odist, size = _geom(n, geom)
    cdt = np.complex128 if prec == "f64" else np.complex64
    x = O.uniform_complex((size,), seed=n % 9973 + len(geom)).astype(cdt)
    n1, n2 = long_split(n)
    print(f"rows two-pass {prec} n={n} ({n1}x{n2}) {geom}", flush=True)
    owned = np.zeros(size, bool)
    for o in range(outer):
        for b in range(batch):
            owned[o * odist + b * dist + np.arange(n) * stride] = True
    for passes in ("2", "4"):
        monkeypatch.setenv("PFX_FOURSTEP_PASSES", passes)
        plan = P.Plan.c2c(prec, n, batch, stride, dist, outer, odist)
        assert plan.describe()[0].endswith(f"fourstep_n{n}_{n1}x{n2}"), plan.describe()
        for direction, sign, scale in ((0, -1, 1.0), (1, +1, 1.0 / n)):
            ref = _oracle(x, n, batch, stride, dist, outer, odist, sign) * scale
            d = to_dev(x)
            plan.execute(direction, d, scale=scale)
            got = to_host(d)
            assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8))
            err = rel_l2(got[owned], ref[owned])
            assert err <= tol(prec, n), (passes, direction, err)
            del ref, got, d
        plan.destroy()


def test_rows_two_pass_has_no_sub_plans_to_tune():
    """Measure planning of a two-pass row plan records nothing (its kernels
    are fixed); the four-pass form records its two tiled sub-plans."""
    import os
    import parafaft_amd as P
    from parafaft_amd import backend as B
    n, batch = 1 << 15, 2
    scratch = to_dev(np.zeros(batch * n, np.complex64))
    for passes, lines in (("2", 0), ("4", 2)):
        os.environ["PFX_FOURSTEP_PASSES"] = passes
        try:
            B.tune_cache_clear()
            plan = P.Plan.c2c("f32", n, batch, 1, n)
            plan.tune(1, scratch)
            assert B.tune_cache_size() == lines, (passes, B.tune_cache_size())
        finally:
            del os.environ["PFX_FOURSTEP_PASSES"]
            B.tune_cache_clear()


def test_rows_two_pass_real_and_chirp():
    """Plans built on long C2C rows take the two-pass form too: FP32 real
    rows of 65536 (half 32768) and a chirp length whose convolution length
    is 2^15 or more (FP32 10007: M = 20480 = 128 x 160, four passes; 16381:
    M = 32768, two passes)."""
    import parafaft_amd as P
    n, batch = 1 << 16, 2
    h = n // 2 + 1
    pd = 2 * h
    real = O.uniform(batch * n, seed=5).reshape(batch, n).astype(np.float32)
    pad = np.zeros((batch, pd), np.float32)
    pad[:, :n] = real
    ref = pad.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, pd)
    r2c = P.Plan.r2c("f32", n, batch, 1, pd)
    c2r = P.Plan.c2r("f32", n, batch, 1, pd)
    d = to_dev(pad)
    r2c.execute(0, d)
    assert rel_l2(to_host(d).copy().view(np.complex64), ref.view(np.complex128)) <= tol("f32", n)
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real.astype(np.float64)) <= 2 * tol("f32", n)
    for m in (10007, 16381):
        x = O.uniform_complex((batch * m,), seed=m).astype(np.complex64)
        plan = P.Plan.c2c("f32", m, batch, 1, m)
        assert "bluestein" in plan.describe()[0]
        r = x.astype(np.complex128).copy()
        O.c2c_batched(r, m, batch, 1, m, -1)
        dd = to_dev(x)
        plan.execute(0, dd)
        assert rel_l2(to_host(dd), r) <= tol("f32", m), m
