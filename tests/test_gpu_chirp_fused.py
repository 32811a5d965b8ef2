"""The chirp convolution (Bluestein) in one kernel (pfx_chirp.inc) for
lengths whose convolution length M has a tiled radix plan (n up to 2048),
and M = 8192 in FP32 (n up to 4096):
premultiply, M-point FFT, filter spectrum, inverse FFT and postmultiply on
one LDS tile.  Against the oracle in rows, batch-contiguous columns, strided
rows with gaps and two-level batches, both directions with a fused scale,
elements the plan does not own bit-identical; the five-pass form
(PFX_CHIRP_FUSED=0) on the same inputs; real lengths whose C2C runs the
chirp (odd lengths, and even ones with a prime half)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import chirp_length, long_split, rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

LENGTHS = [7, 11, 13, 17, 31, 97, 127, 257, 509, 1009, 1021, 2039, 2047, 1001, 1100, 2053, 4093]
GEOMS = {  # batch, stride, dist(n), outer, outer gap
    "rows": (5, 1, lambda n: n, 1, 0),
    "cols": (37, 37, lambda n: 1, 1, 0),
    "gapped": (3, 2, lambda n: 2 * n + 3, 2, 11),
}


def _cases():
    out = []
    for i, n in enumerate(LENGTHS):
        for g in GEOMS:
            out.append(("f64" if (i + len(g)) % 3 == 0 else "f32", n, g))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: "-".join(map(str, c)))
def test_chirp_fused_vs_oracle(case, monkeypatch):
    import parafaft_amd as P
    prec, n, geom = case
    assert long_split(n) is None
    batch, stride, dist, outer, gap = GEOMS[geom]
    dist = dist(n)
    span = (batch - 1) * dist + (n - 1) * stride + 1
    odist = span + gap
    size = (outer - 1) * odist + span
    cdt = np.complex128 if prec == "f64" else np.complex64
    x = O.uniform_complex((size,), seed=n + len(geom)).astype(cdt)
    owned = np.zeros(size, bool)
    for o in range(outer):
        for b in range(batch):
            owned[o * odist + b * dist + np.arange(n) * stride] = True
    M = chirp_length(prec, n)
    print(f"chirp {prec} n={n} M={M} {geom}", flush=True)
    for fused in ("1", "0"):
        monkeypatch.setenv("PFX_CHIRP_FUSED", fused)
        plan = P.Plan.c2c(prec, n, batch, stride, dist, outer, odist)
        assert plan.describe()[0].endswith(f"bluestein_n{n}_m{M}"), plan.describe()
        for direction, sign, scale in ((0, -1, 1.0), (1, +1, 1.0 / n)):
            ref = x.astype(np.complex128).copy()
            for o in range(outer):
                v = ref[o * odist:].copy()
                O.c2c_batched(v, n, batch, stride, dist, sign)
                ref[o * odist:] = v
            ref *= scale
            d = to_dev(x)
            plan.execute(direction, d, scale=scale)
            got = to_host(d)
            assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8))
            err = rel_l2(got[owned], ref[owned])
            assert err <= tol(prec, n), (fused, direction, err)
        plan.destroy()


@pytest.mark.parametrize("prec,n", [("f32", 1001), ("f64", 2 * 1009), ("f32", 2 * 2039),
                                    ("f64", 97)])
def test_chirp_fused_real(prec, n):
    """R2C / C2R whose complex transform is a chirp plan: odd n runs the
    whole row as C2C, even n its half (1009, 2039: primes)."""
    import parafaft_amd as P
    batch = 4
    h = n // 2 + 1
    pd = 2 * h
    rdt = np.float32 if prec == "f32" else np.float64
    real = O.uniform(batch * n, seed=n).reshape(batch, n).astype(rdt)
    pad = np.zeros((batch, pd), rdt)
    pad[:, :n] = real
    ref = pad.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, pd)
    r2c = P.Plan.r2c(prec, n, batch, 1, pd)
    c2r = P.Plan.c2r(prec, n, batch, 1, pd)
    cdt = np.complex64 if prec == "f32" else np.complex128
    d = to_dev(pad)
    r2c.execute(0, d)
    assert rel_l2(to_host(d).copy().view(cdt), ref.view(np.complex128)) <= tol(prec, n)
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real.astype(np.float64)) <= 2 * tol(prec, n)


def test_chirp_fused_empty_and_single():
    """A batch of one transform and a plan over zero transforms."""
    import parafaft_amd as P
    n = 509
    x = O.uniform_complex((n,), seed=4).astype(np.complex64)
    plan = P.Plan.c2c("f32", n, 1, 1, n)
    d = to_dev(x)
    plan.execute(0, d)
    assert rel_l2(to_host(d), O.fftn(x.astype(np.complex128))) <= tol("f32", n)
    empty = P.Plan.c2c("f32", n, 0, 1, n)
    empty.execute(0, d)
