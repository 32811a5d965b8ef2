"""Seeded random geometries against the oracle: C2C lengths that hit tiled
plans, mixed radices and the direct-DFT fallback; row, column and general
strided layouts with gaps between elements; both directions, a fused scale
on some; FP32 and FP64.  Elements a plan does not own must come back
bit-identical (new-array semantics, fft_backend.hpp:107-111, on any
element-aligned layout).  R2C/C2R with padded rows wider than 2(n/2+1)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import c2c_supported, real_supported, rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

_SMOOTH = sorted({a * b for a in (1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 48, 64)
                  for b in (1, 2, 3, 4, 5, 8, 16, 32, 64)} - {1})


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        r = rng.random()
        if r < 0.55:
            n = int(rng.choice(_SMOOTH))
        elif r < 0.8:
            n = int(rng.choice([256, 512, 1024, 2048, 4096]))
        else:
            n = int(rng.integers(2, 400))  # often no tiled plan: the fallback
        batch = int(rng.integers(1, 48 if n <= 1024 else 12))
        layout = ["row", "col", "gen_row", "gen_col"][int(rng.integers(0, 4))]
        prec = "f64" if rng.random() < 0.35 else "f32"
        direction = int(rng.integers(0, 2))
        scale = 1.0 if rng.random() < 0.7 else 1.0 / n
        pad = int(rng.integers(0, 4))
        if layout == "row":
            stride, dist = 1, n + pad
        elif layout == "col":
            stride, dist = batch + pad, 1
        elif layout == "gen_row":
            stride = int(rng.integers(2, 4))
            dist = (n - 1) * stride + 1 + pad
        else:
            dist = int(rng.integers(2, 4))
            stride = (batch - 1) * dist + 1 + pad
        out.append((i, prec, n, batch, stride, dist, direction, scale))
    return out


CASES = _cases(2024, 72)


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c[:6])))
def test_random_c2c_geometry(case):
    import parafaft_amd as P
    i, prec, n, batch, stride, dist, direction, scale = case
    cdt = np.complex64 if prec == "f32" else np.complex128
    size = (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size,), seed=7000 + i).astype(cdt)
    owned = np.zeros(size, bool)
    idx = (np.arange(batch)[:, None] * dist + np.arange(n)[None, :] * stride).ravel()
    owned[idx] = True
    assert owned.sum() == n * batch  # the layout does not overlap itself
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, stride, dist, -1 if direction == 0 else +1)
    ref *= scale
    if not c2c_supported(prec, n):   # no tiled plan and past the direct-DFT limit
        with pytest.raises(P.PfxError):
            P.Plan.c2c(prec, n, batch, stride, dist)
        return
    plan = P.Plan.c2c(prec, n, batch, stride, dist)
    d = to_dev(x)
    plan.execute(direction, d, scale=scale)
    got = to_host(d)
    assert rel_l2(got[owned], ref[owned]) <= tol(prec, n), plan.describe()
    assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8)), plan.describe()


def _real_cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(_SMOOTH + [7, 9, 15, 33, 99]))
        batch = int(rng.integers(1, 24))
        dist = 2 * (n // 2 + 1) + 2 * int(rng.integers(0, 3))
        prec = "f64" if rng.random() < 0.35 else "f32"
        out.append((i, prec, n, batch, dist))
    return out


@pytest.mark.parametrize("case", _real_cases(4048, 30), ids=lambda c: "-".join(map(str, c)))
def test_random_real_geometry(case):
    import parafaft_amd as P
    i, prec, n, batch, dist = case
    rdt = np.float32 if prec == "f32" else np.float64
    h = n // 2 + 1
    buf = O.uniform(batch * dist, seed=9000 + i).reshape(batch, dist).astype(rdt)
    ref = buf.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, dist)
    if not real_supported(prec, n):
        with pytest.raises(P.PfxError):
            P.Plan.r2c(prec, n, batch, 1, dist)
        return
    r2c = P.Plan.r2c(prec, n, batch, 1, dist)
    c2r = P.Plan.c2r(prec, n, batch, 1, dist)
    d = to_dev(buf)
    r2c.execute(0, d)
    got = to_host(d)
    cdt = np.complex64 if prec == "f32" else np.complex128
    assert rel_l2(got[:, :2 * h].copy().view(cdt), ref[:, :2 * h].copy().view(np.complex128)) \
        <= tol(prec, n), r2c.describe()
    # the padding beyond 2(n/2+1) scalars is not touched
    assert np.array_equal(got[:, 2 * h:], buf[:, 2 * h:])
    # C2R of the oracle's spectrum, against the oracle's C2R
    ref2 = ref.copy()
    O.c2r_inplace(ref2, n, batch, 1, dist)
    d2 = to_dev(ref.astype(rdt))
    c2r.execute(1, d2)
    got2 = to_host(d2)
    assert rel_l2(got2[:, :n], ref2[:, :n]) <= tol(prec, n), c2r.describe()
