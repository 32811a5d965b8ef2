"""Radix-32 split-precision kernels (FP16 hi/lo operands on
v_mfma_f32_32x32x16, N = 1024 = 32 x 32 over tiles of 32 batch-contiguous
transforms): every such candidate against the FP64 oracle at the FP32 bound
rel-L2 <= 1e-6 * log2(N), on geometries where tiles straddle outer slices,
on zero, tiny and wide-range butterflies, and for bias over many transforms."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

MODE = "f32_split_f16"
N = 1024


def _split32_candidates(plan):
    idx = []
    for i in range(plan.num_candidates()):
        plan.use_candidate(i)
        if "split32" in plan.describe()[0]:
            idx.append(i)
    assert idx, "no radix-32 split kernel among the candidates"
    return idx


def _ref_two_level(x, n, inner, outer, sign):
    """Oracle: outer slices of `inner` batch-contiguous transforms of length n
    (stride inner), slice o at o * n * inner."""
    ref = x.astype(np.complex128).copy()
    for o in range(outer):
        view = ref[o * n * inner:(o + 1) * n * inner].copy()
        O.c2c_batched(view, n, inner, inner, 1, sign)
        ref[o * n * inner:(o + 1) * n * inner] = view
    return ref


@pytest.mark.parametrize("inner,outer", [(40, 3), (32, 2), (7, 5), (2, 3)])
def test_split32_two_level(inner, outer):
    """The middle axis of an outer x 1024 x inner array in one launch: tiles of
    32 transforms cross slice boundaries unless 32 divides inner; partial last
    tile; both directions with a fused scale."""
    import parafaft_amd as P
    total = inner * outer
    x = O.uniform_complex((N * total,), seed=inner * 31 + outer).astype(np.complex64)
    plan = P.Plan.c2c(MODE, N, inner, inner, 1, outer, N * inner)
    for i in _split32_candidates(plan):
        plan.use_candidate(i)
        name = plan.describe()[0]
        for direction, sign, scale in ((0, -1, 1.0), (1, +1, 1.0 / N)):
            d = to_dev(x)
            plan.execute(direction, d, scale=scale)
            ref = _ref_two_level(x, N, inner, outer, sign) * scale
            err = rel_l2(to_host(d), ref)
            assert err <= tol("f32", N), (name, direction, err)


def test_split32_zero_tiny_and_wide_range():
    """All-zero transforms (the scale's zero case), transforms 1e-30 in size,
    half-zero ones, and magnitudes spanning 2^-20 .. 2^20 inside one
    transform, in one tile."""
    import parafaft_amd as P
    batch = 40
    rng = np.random.default_rng(12)
    x = rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N))
    x[0] = 0.0
    x[1] *= 1e-30
    x[2, ::2] = 0.0
    x[3] *= 2.0 ** rng.uniform(-20, 20, N)
    x = x.astype(np.complex64)
    flat = np.ascontiguousarray(x.T).reshape(-1)
    plan = P.Plan.c2c(MODE, N, batch, batch, 1)
    ref = flat.astype(np.complex128).copy()
    O.c2c_batched(ref, N, batch, batch, 1, -1)
    r2 = ref.reshape(N, batch).T
    for i in _split32_candidates(plan):
        plan.use_candidate(i)
        name = plan.describe()[0]
        d = to_dev(flat)
        plan.execute(0, d)
        got = to_host(d)
        assert np.all(np.isfinite(got)), name
        g2 = got.reshape(N, batch).T
        assert rel_l2(got, ref) <= tol("f32", N), name
        assert np.all(g2[0] == 0), name
        for t in (1, 3):  # accuracy relative to the transform itself
            assert rel_l2(g2[t], r2[t]) <= tol("f32", N), (name, t)


def test_split32_unbiased():
    """Mean error of every output bin over 8192 transforms stays within a few
    standard errors of zero (the Thue-Morse signs cancel the matrix cores'
    accumulation bias in both passes)."""
    import parafaft_amd as P
    batch = 1 << 13
    x = O.uniform_complex((N * batch,), seed=23).astype(np.complex64)
    plan = P.Plan.c2c(MODE, N, batch, batch, 1)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, N, batch, batch, 1, -1)
    for i in _split32_candidates(plan):
        plan.use_candidate(i)
        d = to_dev(x)
        plan.execute(0, d)
        e = (to_host(d).astype(np.complex128) - ref).reshape(N, batch).T
        for part in (e.real, e.imag):
            z = part.mean(axis=0) / (part.std(axis=0) / np.sqrt(batch))
            assert abs(z.mean()) < 1.0, (plan.describe()[0], z.mean())
            assert np.abs(z).max() < 8.0, (plan.describe()[0], np.abs(z).max())


def test_split32_is_a_tuning_candidate_and_bf16_is_not():
    """The Measure tuner sees the radix-32 kernels for FP16 column plans; the
    BF16 form has none for columns (its operand splits do not fit the
    registers of the 32-column tile)."""
    import parafaft_amd as P
    p16 = P.Plan.c2c("f32_split_f16", N, 64, 64, 1)
    _split32_candidates(p16)
    pbf = P.Plan.c2c("f32_split_bf16", N, 64, 64, 1)
    names = []
    for i in range(pbf.num_candidates()):
        pbf.use_candidate(i)
        names.append(pbf.describe()[0])
    assert not any("split32" in nm for nm in names), names
