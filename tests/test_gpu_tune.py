"""Measure/Patient planning (pfx_plan_tune) and the tuning cache (the FFTW
wisdom analogue, backend/fftw3/fft_backend_fftw.hpp:666-694)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu


def test_tune_picks_a_correct_kernel_and_caches(tmp_path):
    import parafaft_amd as P
    from parafaft_amd import backend as B
    B.tune_cache_clear()
    n, batch = 1024, 4096
    x = O.uniform_complex((n * batch,), seed=17).astype(np.complex64)
    scratch = to_dev(np.zeros_like(x))
    plan = P.Plan.c2c("f32", n, batch, batch, 1)
    default_name = plan.describe()[0]
    plan.tune(1, scratch)
    tuned = plan.describe()[0]
    assert B.tune_cache_size() == 1
    d = to_dev(x)
    plan.execute(0, d)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, batch, batch, 1, -1)
    assert rel_l2(to_host(d), ref) <= tol("f32", n), (default_name, tuned)
    # persist, clear, reload: a new plan of the same geometry starts tuned
    path = str(tmp_path / "pfx_tune.txt")
    B.tune_cache_save(path)
    assert open(path).read().count("\n") >= 2
    B.tune_cache_clear()
    assert B.tune_cache_size() == 0
    B.tune_cache_load(path)
    assert B.tune_cache_size() == 1
    again = P.Plan.c2c("f32", n, batch, batch, 1)
    assert again.describe()[0] == tuned
    B.tune_cache_clear()


def test_backend_measure_flag_through_driver():
    import parafaft_amd as P
    from parafaft_amd import backend as B
    B.tune_cache_clear()
    n = 64
    x = O.uniform_complex((n, n, n), seed=2).astype(np.complex64)
    fft = P.ParaFaFT_C2C([n] * 3, precision="f32", plan_flag=P.FFTPlanFlag.Patient)
    assert B.tune_cache_size() == 3
    d = to_dev(x)
    fft.forward(d)
    assert rel_l2(to_host(d), O.fftn(x.astype(np.complex128))) <= tol("f32", n ** 3)
    B.tune_cache_clear()


GEOMS = [  # (prec, n, batch, stride, dist, outer, outer_dist)
    ("f32", 1024, 100, 1, 1024, 1, 0), ("f32", 1024, 75, 75, 1, 1, 0),
    ("f32", 1024, 40, 40, 1, 3, 1024 * 40), ("f32", 2048, 50, 1, 2048, 1, 0),
    ("f32", 2048, 33, 33, 1, 1, 0), ("f32", 4096, 20, 20, 1, 1, 0), ("f32", 512, 70, 70, 1, 1, 0),
    ("f32", 256, 129, 129, 1, 1, 0), ("f64", 1024, 40, 40, 1, 1, 0), ("f64", 1024, 33, 1, 1024, 1, 0),
    # short batch-contiguous lengths (the wide 128- and 256-column tiles of
    # 64 and 128 points; partial last tiles; a two-level batch)
    ("f32", 128, 300, 300, 1, 1, 0), ("f32", 128, 130, 130, 1, 2, 128 * 130),
    ("f32", 64, 520, 520, 1, 1, 0),
    ("f32_split_f16", 1024, 50, 50, 1, 1, 0), ("f32_split_bf16", 1024, 50, 1, 1024, 1, 0),
]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "-".join(map(str, g[:5])))
def test_every_candidate_kernel(geom):
    """Every kernel the tuner may pick for a geometry gives the oracle's result."""
    import parafaft_amd as P
    prec, n, batch, stride, dist, outer, odist = geom
    cdt = np.complex128 if prec == "f64" else np.complex64
    size = (outer - 1) * odist + (batch - 1) * dist + (n - 1) * stride + 1
    x = O.uniform_complex((size,), seed=n + batch).astype(cdt)
    ref = x.astype(np.complex128).copy()
    for o in range(outer):
        seg = np.ascontiguousarray(ref[o * odist:])
        O.c2c_batched(seg, n, batch, stride, dist, -1)
        ref[o * odist:] = seg
    plan = P.Plan.c2c(prec, n, batch, stride, dist, outer, odist)
    k = plan.num_candidates()
    assert k >= 1
    names = set()
    for i in range(k):
        plan.use_candidate(i)
        name = plan.describe()[0]
        names.add(name)
        d = to_dev(x)
        plan.execute(0, d)
        assert rel_l2(to_host(d), ref) <= tol("f64" if prec == "f64" else "f32", n), name
    assert len(names) == k


def test_cache_line_with_a_foreign_kernel_is_ignored(tmp_path):
    """A cache file is input: a line that pairs a geometry with a kernel of
    another length leaves the plan on its own default kernel (which then
    gives the oracle's result)."""
    import parafaft_amd as P
    from parafaft_amd import backend as B
    B.tune_cache_clear()
    geoms = [(1024, 96), (2048, 48)]
    defaults = {}
    for n, batch in geoms:
        scratch = to_dev(np.zeros(n * batch, np.complex64))
        plan = P.Plan.c2c("f32", n, batch, batch, 1)
        defaults[n] = plan.describe()[0]
        plan.tune(1, scratch)
    path = str(tmp_path / "pfx_tune.txt")
    B.tune_cache_save(path)
    lines = [l for l in open(path).read().splitlines() if l and not l.startswith("#")]
    assert len(lines) == 2
    (k1, v1), (k2, v2) = [l.split("\t") for l in lines]
    with open(path, "w") as f:
        f.write(f"{k1}\t{v2}\n{k2}\t{v1}\n")
    B.tune_cache_clear()
    assert B.tune_cache_load(path) == 0
    for n, batch in geoms:
        plan = P.Plan.c2c("f32", n, batch, batch, 1)
        assert plan.describe()[0] == defaults[n]
        x = O.uniform_complex((n * batch,), seed=n).astype(np.complex64)
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, batch, 1, -1)
        d = to_dev(x)
        plan.execute(0, d)
        assert rel_l2(to_host(d), ref) <= tol("f32", n)
        # a Measure tune must not take the foreign line either: it measures
        # afresh and keeps a kernel of this plan's own length
        scratch = to_dev(np.zeros(n * batch, np.complex64))
        plan.tune(1, scratch)
        assert f"_n{n}_" in plan.describe()[0], plan.describe()
        d = to_dev(x)
        plan.execute(0, d)
        assert rel_l2(to_host(d), ref) <= tol("f32", n)
    B.tune_cache_clear()


def test_cache_load_counts_dropped_lines(tmp_path):
    """Lines whose key is not in the cache's format, or whose kernel this
    build lacks, are dropped and counted; valid lines still load."""
    import parafaft_amd as P
    from parafaft_amd import backend as B
    B.tune_cache_clear()
    scratch = to_dev(np.zeros(256 * 64, np.complex64))
    P.Plan.c2c("f32", 256, 64, 64, 1).tune(1, scratch)
    path = str(tmp_path / "pfx_tune.txt")
    B.tune_cache_save(path)
    good = [l for l in open(path).read().splitlines() if l and not l.startswith("#")]
    assert len(good) == 1
    key, val = good[0].split("\t")
    with open(path, "a") as f:
        f.write(f"not a key\t{val}\n")                 # bad key
        f.write(f"{key.replace('n256', 'n512')}\t0|0|999|1|1|nope|order=1\n")  # no such kernel
        f.write("no tab on this line\n")
    B.tune_cache_clear()
    assert B.tune_cache_load(path) == 3
    assert B.tune_cache_size() == 1
    B.tune_cache_clear()
