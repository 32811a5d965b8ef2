"""Rows beyond 4096 points in one kernel: 8192 and 16384 (FP32 and FP64)
in the row layout run register-resident Stockham kernels (twiddles
from the global table, a one-component LDS exchange); batch-contiguous
layouts of these lengths keep the four-step plan.  Every candidate kernel on
contiguous, gapped, strided and two-level row geometries, both directions
with a fused scale, against the oracle; elements the plan does not own stay
untouched."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, tol

pytestmark = pytest.mark.gpu

ROWS = [("f32", 8192), ("f64", 8192), ("f32", 16384), ("f64", 16384)]
GEOMS = {  # batch, stride, dist, outer, outer gap
    "contiguous": (5, 1, None, 1, 0),
    "gapped": (3, 1, 7, 2, 33),
    "strided": (2, 3, 1, 1, 0),
}


def _geom(n, name):
    batch, stride, dist, outer, gap = GEOMS[name]
    dist = n * stride + (dist or 0) if name != "contiguous" else n
    if name == "strided":
        dist = 3 * (n - 1) + 1 + 5
    span = (batch - 1) * dist + (n - 1) * stride + 1
    outer_dist = span + gap
    size = (outer - 1) * outer_dist + span
    return batch, stride, dist, outer, outer_dist, size


@pytest.mark.parametrize("prec,n", ROWS)
@pytest.mark.parametrize("geom", list(GEOMS))
def test_long_rows_every_candidate(prec, n, geom):
    import parafaft_amd as P
    batch, stride, dist, outer, outer_dist, size = _geom(n, geom)
    cdt = np.complex128 if prec == "f64" else np.complex64
    x = O.uniform_complex((size,), seed=n + len(geom)).astype(cdt)
    refs = {}
    for direction, sign in ((0, -1), (1, +1)):
        r = x.astype(np.complex128).copy()
        for o in range(outer):
            v = r[o * outer_dist:].copy()
            O.c2c_batched(v, n, batch, stride, dist, sign)
            r[o * outer_dist:] = v
        refs[direction] = r
    owned = refs[0] != x
    plan = P.Plan.c2c(prec, n, batch, stride, dist, outer, outer_dist)
    assert plan.num_candidates() >= 1
    names = []
    for i in range(plan.num_candidates()):
        plan.use_candidate(i)
        name = plan.describe()[0]
        names.append(name)
        assert f"_row_n{n}_" in name, name
        for direction, scale in ((0, 1.0), (1, 1.0 / n)):
            d = to_dev(x)
            plan.execute(direction, d, scale=scale)
            got = to_host(d)
            assert np.array_equal(got[~owned].view(np.uint8), x[~owned].view(np.uint8)), name
            err = rel_l2(got[owned], refs[direction][owned] * scale)
            assert err <= tol(prec, n), (name, direction, err)
    assert len(set(names)) == len(names)


@pytest.mark.parametrize("prec,n", ROWS)
def test_long_rows_plan_kinds(prec, n):
    """Rows run one tiled kernel; batch-contiguous columns of the same length
    run the four-step plan; the tuner picks among the row kernels."""
    import parafaft_amd as P
    row = P.Plan.c2c(prec, n, 4, 1, n)
    assert "fourstep" not in row.describe()[0] and f"_row_n{n}_" in row.describe()[0]
    col = P.Plan.c2c(prec, n, 4, 4, 1)
    assert "fourstep" in col.describe()[0], col.describe()
    cdt = np.complex128 if prec == "f64" else np.complex64
    row.tune(1, to_dev(np.zeros(4 * n, cdt)))
    assert f"_row_n{n}_" in row.describe()[0]
    x = O.uniform_complex((4 * n,), seed=3).astype(cdt)
    ref = x.astype(np.complex128).copy()
    O.c2c_batched(ref, n, 4, 1, n, -1)
    d = to_dev(x)
    row.execute(0, d)
    assert rel_l2(to_host(d), ref) <= tol(prec, n)


@pytest.mark.parametrize("prec,n", [("f32", 16384), ("f64", 16384), ("f32", 32768),
                                    ("f64", 32768)])
def test_long_real_rows_through_one_kernel_half(prec, n):
    """Even real lengths past 8192: FP32 16384 runs one R2C kernel; FP64
    16384 and FP32 / FP64 32768 run their half (8192, 16384) as a one-kernel
    C2C row, then the R2C post-pass; C2R back."""
    import parafaft_amd as P
    batch = 3
    h = n // 2 + 1
    pd = 2 * h
    rdt = np.float32 if prec == "f32" else np.float64
    real = O.uniform(batch * n, seed=n + 2).reshape(batch, n).astype(rdt)
    pad = np.zeros((batch, pd), dtype=rdt)
    pad[:, :n] = real
    ref = pad.astype(np.float64).copy()
    O.r2c_inplace(ref, n, batch, 1, pd)
    r2c = P.Plan.r2c(prec, n, batch, 1, pd)
    c2r = P.Plan.c2r(prec, n, batch, 1, pd)
    d = to_dev(pad)
    r2c.execute(0, d)
    cdt = np.complex64 if prec == "f32" else np.complex128
    got = to_host(d).copy().view(cdt)
    assert rel_l2(got, ref.view(np.complex128)) <= tol(prec, n)
    c2r.execute(1, d, scale=1.0 / n)
    assert rel_l2(to_host(d)[:, :n], real.astype(np.float64)) <= 2 * tol(prec, n)


@pytest.mark.parametrize("prec,n", [("f32", 4099), ("f32", 8191), ("f64", 3001)])
def test_chirp_with_one_kernel_convolution(prec, n):
    """Bluestein lengths whose power-of-two convolution length M is 8192 or
    16384: the M-point transforms now run one kernel."""
    import parafaft_amd as P
    batch = 3
    cdt = np.complex128 if prec == "f64" else np.complex64
    x = O.uniform_complex((batch * n,), seed=n).astype(cdt)
    plan = P.Plan.c2c(prec, n, batch, 1, n)
    assert "bluestein" in plan.describe()[0]
    for direction, sign in ((0, -1), (1, +1)):
        ref = x.astype(np.complex128).copy()
        O.c2c_batched(ref, n, batch, 1, n, sign)
        d = to_dev(x)
        plan.execute(direction, d)
        assert rel_l2(to_host(d), ref) <= tol(prec, n)


@pytest.mark.parametrize("kind", ["r2c", "c2r"])
def test_real_16384_every_candidate(kind):
    """Every R2C / C2R kernel of real length 16384 (FP32, half 8192 in one
    kernel) on gapped padded rows, against the oracle; the padding and the
    gaps stay untouched by R2C."""
    import parafaft_amd as P
    n, batch = 16384, 5
    h = n // 2 + 1
    pd = 2 * h + 6
    real = O.uniform(batch * n, seed=17).reshape(batch, n)
    pad = np.zeros((batch, pd))
    pad[:, :n] = real
    spec = pad.copy()
    O.r2c_inplace(spec, n, batch, 1, pd)
    if kind == "r2c":
        src, ref = pad.astype(np.float32), spec
        plan = P.Plan.r2c("f32", n, batch, 1, pd)
    else:
        back = spec.copy()
        O.c2r_inplace(back, n, batch, 1, pd)
        src, ref = spec.astype(np.float32), back
        plan = P.Plan.c2r("f32", n, batch, 1, pd)
    assert plan.num_candidates() >= 3
    for i in range(plan.num_candidates()):
        plan.use_candidate(i)
        name = plan.describe()[0]
        assert f"{kind}_row_n8192_" in name, name
        d = to_dev(src)
        plan.execute(0 if kind == "r2c" else 1, d)
        got = to_host(d)
        if kind == "r2c":
            assert rel_l2(got[:, :2 * h].copy().view(np.complex64),
                          ref[:, :2 * h].copy().view(np.complex128)) <= tol("f32", n), name
            assert np.all(got[:, 2 * h:] == 0), name
        else:
            assert rel_l2(got[:, :n], ref[:, :n]) <= tol("f32", n), name


@pytest.mark.parametrize("case", [
    ("c2c", "f32", 32768, 2, 1, 32768),   # four-step rows (four passes)
    ("c2c", "f32", 8192, 6, 6, 1),        # four-step columns (fused, two passes)
    ("c2c", "f64", 8192, 3, 1, 8192),     # one-kernel rows
    ("c2c", "f32", 8191, 3, 1, 8191),     # chirp convolution
    ("r2c", "f32", 32768, 2, 1, 32770),   # real: half C2C + post-pass
])
def test_long_plans_under_graph_capture(case):
    """A long plan's execution captured into a HIP graph and replayed several
    times, with eager work and synchronisations between the replays: every
    replay equals the eager result bit for bit (the plan's scratch and tables
    are allocated at creation; no allocation or memset is captured)."""
    import parafaft_amd as P
    from parafaft_amd.batch import Graph
    from tests.gpu_util import torch
    t = torch()
    kind, prec, n, batch, stride, dist = case
    stream = t.cuda.Stream()
    if kind == "c2c":
        cdt = np.complex128 if prec == "f64" else np.complex64
        size = (batch - 1) * dist + (n - 1) * stride + 1
        host = O.uniform_complex((size,), seed=n + batch).astype(cdt)
        plan = P.Plan.c2c(prec, n, batch, stride, dist)
        direction = 0
    else:
        host = np.zeros((batch, dist), np.float32)
        host[:, :n] = O.uniform(batch * n, seed=n).reshape(batch, n)
        host = host.reshape(-1)
        plan = P.Plan.r2c(prec, n, batch, 1, dist)
        direction = 0
    d = to_dev(host)
    plan.execute(direction, d, stream=stream.cuda_stream)
    stream.synchronize()
    eager = to_host(d)
    g = Graph(stream.cuda_stream, lambda: plan.execute(direction, d, stream=stream.cuda_stream))
    for rep in range(4):
        d.copy_(t.from_numpy(host).to("cuda"))
        t.cuda.synchronize()
        _ = (t.ones(1000, device="cuda") * 2).sum().item()  # eager work between replays
        g.launch(stream.cuda_stream)
        stream.synchronize()
        assert np.array_equal(to_host(d).view(np.uint8), eager.view(np.uint8)), (case, rep)


@pytest.mark.parametrize("kind,prec,shape", [
    ("c2c", "f32", (3, 8192, 5)),     # long middle axis: two-level batch-contiguous four-step
    ("c2c", "f64", (8192, 6)),        # long outer axis, batch-contiguous
    ("c2c", "f32", (2, 5, 16384)),    # long rows in one kernel
    ("c2c", "f32", (32768, 3)),       # long outer axis beyond one kernel
    ("r2c", "f32", (4, 8192, 10)),    # R2C with a long middle C2C axis
    ("r2c", "f64", (3, 16384)),       # long real rows (FP64: half 8192 as one C2C kernel)
])
@pytest.mark.parametrize("fast", [True, False])
def test_long_axes_in_nd_arrays(kind, prec, shape, fast):
    """Long axes anywhere in an N-D array through both one-rank driver modes:
    forward against the oracle, backward with the fused 1/N scale back to the
    input."""
    import parafaft_amd as P
    shape = list(shape)
    N = int(np.prod(shape))
    t = tol(prec, N)
    if kind == "c2c":
        cdt = np.complex64 if prec == "f32" else np.complex128
        x = O.uniform_complex(tuple(shape), seed=N % 1000).astype(cdt)
        ref = O.fftn(x.astype(np.complex128))
        fft = P.ParaFaFT_C2C(shape, precision=prec, fast=fast)
        d = to_dev(x)
        fft.forward(d)
        assert rel_l2(to_host(d), ref) <= t
        fft.backward(d, scale=1.0 / N)
        assert rel_l2(to_host(d), x) <= 2 * t
    else:
        rdt = np.float32 if prec == "f32" else np.float64
        cdt = np.complex64 if prec == "f32" else np.complex128
        x = O.uniform(N, seed=N % 997).reshape(shape).astype(rdt)
        ref = O.rfftn_padded(x.astype(np.float64))
        fft = P.ParaFaFT_R2C(shape, precision=prec, fast=fast)
        h = shape[-1] // 2 + 1
        rows = int(np.prod(shape[:-1]))
        pad = np.zeros(fft.get_required_output_size(), rdt)
        pad[:rows * 2 * h].reshape(rows, 2 * h)[:, :shape[-1]] = x.reshape(rows, -1)
        d = to_dev(pad)
        fft.forward_in_place(d)
        got = to_host(d)[:rows * 2 * h].copy().view(cdt).reshape(shape[:-1] + [h])
        assert rel_l2(got, ref) <= t
        fft.backward_in_place(d, scale=1.0 / N)
        back = to_host(d)[:rows * 2 * h].reshape(rows, 2 * h)[:, :shape[-1]]
        assert rel_l2(back.reshape(shape), x) <= 2 * t
