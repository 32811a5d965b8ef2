"""Split-precision MFMA form: radix-16 sub-DFTs on v_mfma_f32_16x16x16 with
FP16 hi/lo (per-butterfly power-of-two scaling) or BF16 hi/mid/lo operand splits,
FP32 accumulation; every result against the FP64 oracle, at the FP32 bound
rel-L2 <= 1e-6 * log2(N)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

MODES = ["f32_split_f16", "f32_split_bf16"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [256, 1024, 2048, 4096])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_split_1d(mode, n, layout):
    import parafaft_amd as P
    batch = 33
    x = O.uniform_complex((n * batch,), seed=n + len(mode)).astype(np.complex64)
    if layout == "row":
        stride, dist = 1, n
    else:
        stride, dist = batch, 1
    plan = P.Plan.c2c(mode, n, batch, stride, dist)
    name = plan.describe()[0]
    assert "split" in name, name
    for direction, sign in ((0, -1), (1, +1)):
        d = to_dev(x)
        plan.execute(direction, d)
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, stride, dist, sign)
        err = rel_l2(to_host(d), ref)
        assert err <= tol("f32", n), (name, direction, err)


@pytest.mark.parametrize("mode", MODES)
def test_split_wide_dynamic_range(mode):
    """Inputs spanning 2^-20 .. 2^20 within one transform (the FP16 form's
    per-butterfly scaling keeps the large values finite)."""
    import parafaft_amd as P
    n, batch = 1024, 8
    rng = np.random.default_rng(5)
    mag = 2.0 ** rng.uniform(-20, 20, n * batch)
    x = (mag * np.exp(2j * np.pi * rng.random(n * batch))).astype(np.complex64)
    plan = P.Plan.c2c(mode, n, batch, 1, n)
    d = to_dev(x)
    plan.execute(0, d)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, 1, n, -1)
    got = to_host(d)
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= tol("f32", n)


@pytest.mark.parametrize("mode", MODES)
def test_split_2d_through_fast_path(mode):
    import parafaft_amd as P
    n = 1024
    x = O.uniform_complex((256, n), seed=3).astype(np.complex64)
    fft = P.ParaFaFT_C2C([256, n], precision=mode, fast=True)
    d = to_dev(x)
    fft.forward(d)
    ref = O.fftn(x.astype(np.complex128))
    assert rel_l2(to_host(d), ref) <= tol("f32", 256 * n)
    fft.backward(d, scale=1.0 / (256 * n))
    assert rel_l2(to_host(d), x) <= tol("f32", 256 * n)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [256, 512, 1024, 2048])
@pytest.mark.parametrize("layout", ["row", "col"])
def test_split_every_candidate(mode, n, layout):
    """Every split kernel the planner may pick (LDS-staged and register-
    resident variants), both directions, with a fused scale, on a batch that
    leaves partial workgroups."""
    import parafaft_amd as P
    batch = 37
    x = O.uniform_complex((n * batch,), seed=n + 7).astype(np.complex64)
    stride, dist = (1, n) if layout == "row" else (batch, 1)
    plan = P.Plan.c2c(mode, n, batch, stride, dist)
    names = []
    for i in range(plan.num_candidates()):
        plan.use_candidate(i)
        name = plan.describe()[0]
        names.append(name)
        for direction, sign, scale in ((0, -1, 1.0), (1, +1, 1.0 / n)):
            d = to_dev(x)
            plan.execute(direction, d, scale=scale)
            ref = x.astype(np.complex128).copy()
            O.c2c_batched(ref, n, batch, stride, dist, sign)
            err = rel_l2(to_host(d), ref * scale)
            assert err <= tol("f32", n), (name, direction, err)
    # a register-resident kernel is among the candidates for every case
    assert any("split-reg" in nm for nm in names), names


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("layout", ["row", "col"])
def test_split_zero_and_tiny_butterflies(mode, layout):
    """Butterflies whose 16 inputs are all zero (the scale's zero case) or
    far below FP16's normal range sit next to unit-size ones."""
    import parafaft_amd as P
    n, batch = 1024, 32
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n)))
    x[0] = 0.0
    x[1] *= 1e-30
    x[2, ::2] = 0.0
    x = x.astype(np.complex64)
    if layout == "row":
        flat, stride, dist = x.reshape(-1), 1, n
    else:
        flat, stride, dist = np.ascontiguousarray(x.T).reshape(-1), batch, 1
    plan = P.Plan.c2c(mode, n, batch, stride, dist)
    d = to_dev(flat)
    plan.execute(0, d)
    got = to_host(d)
    ref = flat.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, stride, dist, -1)
    assert np.all(np.isfinite(got))
    assert rel_l2(got, ref) <= tol("f32", n)
    g2 = got.reshape(batch, n) if layout == "row" else got.reshape(n, batch).T
    r2 = ref.reshape(batch, n) if layout == "row" else ref.reshape(n, batch).T
    assert np.all(g2[0] == 0)
    # the tiny transform keeps FP32-level accuracy relative to itself
    assert rel_l2(g2[1], r2[1]) <= tol("f32", n)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("layout", ["row", "col"])
def test_split_unbiased(mode, layout):
    """The matrix cores' FP32 accumulation rounds with a bias toward -inf; the
    kernels flip the sign of each butterfly's pass in a Thue-Morse pattern so
    the bias cancels.  Over many transforms the mean error of every output
    bin stays within a few standard errors of zero (with one sign for all
    butterflies it sits near -11 standard errors at this batch, and such a
    bias adds up coherently in the DC sums of a later axis)."""
    import parafaft_amd as P
    n, batch = 256, 1 << 14
    x = O.uniform_complex((n * batch,), seed=21).astype(np.complex64)
    stride, dist = (1, n) if layout == "row" else (batch, 1)
    plan = P.Plan.c2c(mode, n, batch, stride, dist)
    d = to_dev(x)
    plan.execute(0, d)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, stride, dist, -1)
    e = to_host(d).astype(np.complex128) - ref
    e = e.reshape(batch, n) if layout == "row" else e.reshape(n, batch).T
    for part in (e.real, e.imag):
        z = part.mean(axis=0) / (part.std(axis=0) / np.sqrt(batch))
        assert abs(z.mean()) < 1.0, (plan.describe()[0], z.mean())
        assert np.abs(z).max() < 8.0, (plan.describe()[0], np.abs(z).max())


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,batch", [(1 << 20, 2), (8192, 5), (1000, 7)])
@pytest.mark.parametrize("flag", [0, 1])
def test_split_long_lengths(mode, n, batch, flag):
    """Lengths beyond one kernel in a split precision: the four-step plan
    (2^20 = 1024 x 1024, whose first pass runs 1024-point columns out of
    place, 8192 = 64 x 128) and the chirp convolution (1000), with Estimate
    (flag 0) and Measure (flag 1) planning of their sub-plans; rows and
    batch-contiguous layouts, both directions, against the oracle."""
    import parafaft_amd as P
    for layout in ("row", "col"):
        stride, dist = (1, n) if layout == "row" else (batch, 1)
        x = O.uniform_complex((n * batch,), seed=n % 97 + batch).astype(np.complex64)
        plan = P.Plan.c2c(mode, n, batch, stride, dist)
        if flag:
            plan.tune(flag, to_dev(np.zeros_like(x)))
        name = plan.describe()[0]
        for direction, sign in ((0, -1), (1, +1)):
            d = to_dev(x)
            plan.execute(direction, d)
            ref = x.astype(np.complex128).copy()
            O.c2c_batched(ref, n, batch, stride, dist, sign)
            err = rel_l2(to_host(d), ref)
            assert err <= tol("f32", n), (name, layout, direction, err)


def test_split_every_kernel_in_long_plan_sub_passes(tmp_path):
    """A long plan's first tiled pass runs out of place (loads from the
    caller's array, stores into the plan's scratch).  Every split kernel of
    each sub-plan geometry, forced in turn through a tuning-cache line, must
    load from the source: the result matches the oracle."""
    from tests.gpu_util import forced_sub_plan_kernels
    n, batch = 1 << 20, 2
    x = O.uniform_complex((n * batch,), seed=41).astype(np.complex64)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, batch, 1, -1)
    for mode in MODES:
        forced = 0
        for key, name, plan in forced_sub_plan_kernels(mode, n, batch, batch, 1,
                                                       str(tmp_path / "cache.txt")):
            d = to_dev(x)
            plan.execute(0, d)
            err = rel_l2(to_host(d), ref)
            assert err <= tol("f32", n), (mode, key, name, err)
            forced += 1
        assert forced >= 4, forced
