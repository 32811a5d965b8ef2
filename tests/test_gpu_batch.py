"""Grouped execution of many uneven transforms and HIP graph replay."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, torch

pytestmark = pytest.mark.gpu

SMALL = [("c2c", (48, 48, 48)), ("c2c", (64, 64, 64)), ("c2c", (16, 24)), ("c2c", (1024, 32)),
         ("r2c", (20, 24, 16)), ("r2c", (12, 10, 15)), ("c2c", (8, 8, 8, 8)), ("c2c", (96,)),
         ("r2c", (80, 12, 10)), ("c2c", (48, 48, 48)), ("c2c", (7, 11, 13)), ("r2c", (33,)),
         ("c2c", (120, 8, 10)), ("c2c", (32, 32, 32, 32)), ("c2c", (256, 256))]


def _layout(items, cdt):
    offs, pos = [], 0
    from parafaft_amd.batch import storage_elems
    for kind, shape in items:
        offs.append(pos)
        pos += (storage_elems(kind, shape) + 63) // 64 * 64
    return offs, pos


def _inputs(items, seed):
    out = []
    for i, (kind, shape) in enumerate(items):
        if kind == "c2c":
            out.append(O.uniform_complex(shape, seed=seed + i))
        else:
            out.append(O.uniform(int(np.prod(shape)), seed=seed + i).reshape(shape))
    return out


def _pack_inputs(items, xs, offs, total, cdt, rdt):
    buf = np.zeros(total, cdt)
    for (kind, shape), x, o in zip(items, xs, offs):
        if kind == "c2c":
            buf[o:o + x.size] = x.ravel().astype(cdt)
        else:
            n = shape[-1]
            h = n // 2 + 1
            pad = np.zeros(shape[:-1] + (2 * h,), rdt)
            pad[..., :n] = x
            buf[o:o + pad.size // 2] = pad.ravel().view(cdt)
    return buf


def _refs(items, xs):
    refs = []
    for (kind, shape), x in zip(items, xs):
        refs.append(O.fftn(x) if kind == "c2c" else O.rfftn_padded(x))
    return refs


def _check(items, buf, offs, refs, prec, cdt):
    tol = 1e-13 if prec == "f64" else 1e-6 * 22
    worst = 0
    for (kind, shape), o, r in zip(items, offs, refs):
        got = buf[o:o + r.size].reshape(r.shape)
        e = rel_l2(got, r)
        worst = max(worst, e)
        assert e <= tol, (kind, shape, e)
    return worst


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batch_grouped_and_each_vs_oracle(prec):
    from parafaft_amd.batch import Batch
    cdt = np.complex64 if prec == "f32" else np.complex128
    rdt = np.float32 if prec == "f32" else np.float64
    offs, total = _layout(SMALL, cdt)
    xs = _inputs(SMALL, 40)
    host = _pack_inputs(SMALL, xs, offs, total, cdt, rdt)
    refs = _refs(SMALL, [x.astype(cdt if k == "c2c" else rdt).astype(
        np.complex128 if k == "c2c" else np.float64) for (k, _), x in zip(SMALL, xs)])
    for mode in ("grouped", "each"):
        d = to_dev(host)
        esz = np.dtype(cdt).itemsize
        b = Batch(prec, [(k, s, d.data_ptr() + o * esz) for (k, s), o in zip(SMALL, offs)])
        g, e = b.launches(0)
        assert g < e
        (b.execute if mode == "grouped" else b.execute_each)(0)
        _check(SMALL, to_host(d), offs, refs, prec, cdt)
        # backward: C2C gives N x, R2C gives N x in the real slots
        (b.execute if mode == "grouped" else b.execute_each)(1)
        back = to_host(d)
        for (kind, shape), o, x in zip(SMALL, offs, xs):
            N = int(np.prod(shape))
            if kind == "c2c":
                got = back[o:o + x.size].reshape(shape)
            else:
                n = shape[-1]
                h = n // 2 + 1
                got = back[o:o + int(np.prod(shape[:-1])) * h].view(rdt).reshape(
                    shape[:-1] + (2 * h,))[..., :n]
            assert rel_l2(got, N * x) <= (1e-13 if prec == "f64" else 1e-6 * 25), (kind, shape)


def test_graph_replays_match_eager():
    """A captured graph replayed several times, with eager work and syncs in
    between; every replay equals the eager result."""
    import parafaft_amd as P
    from parafaft_amd.batch import Batch, Graph
    t = torch()
    prec, cdt, rdt = "f32", np.complex64, np.float32
    offs, total = _layout(SMALL, cdt)
    xs = _inputs(SMALL, 90)
    host = _pack_inputs(SMALL, xs, offs, total, cdt, rdt)
    d = to_dev(host)
    b = Batch(prec, [(k, s, d.data_ptr() + o * 8) for (k, s), o in zip(SMALL, offs)])
    stream = t.cuda.Stream()
    b.execute(0, stream.cuda_stream)
    t.cuda.synchronize()
    eager = to_host(d)
    g = Graph(stream.cuda_stream, lambda: b.execute(0, stream.cuda_stream))
    for rep in range(4):
        d.copy_(t.from_numpy(host).to("cuda"))
        t.cuda.synchronize()
        _ = (t.ones(1000, device="cuda") * 2).sum().item()   # eager work between replays
        g.launch(stream.cuda_stream)
        stream.synchronize()
        assert np.array_equal(to_host(d), eager), rep
    # a whole-array fast-path plan under capture as well
    n = 64
    x = O.uniform_complex((n, n, n), seed=5).astype(np.complex64)
    fft = P.ParaFaFT_C2C([n] * 3, precision="f32", fast=True, stream=stream.cuda_stream)
    dx = to_dev(x)
    gx = Graph(stream.cuda_stream, lambda: fft._nd.execute(0, dx, 1.0, stream.cuda_stream))
    ref = O.fftn(x.astype(np.complex128))
    for rep in range(3):
        dx.copy_(t.from_numpy(x).to("cuda"))
        t.cuda.synchronize()
        gx.launch(stream.cuda_stream)
        stream.synchronize()
        assert rel_l2(to_host(dx), ref) <= 1e-6 * 18


WORKLOAD_SHAPES = [("r2c", (80, 80, 80)), ("r2c", (80, 80, 80)), ("r2c", (120, 120, 120)),
                   ("r2c", (160, 160, 160)), ("r2c", (120, 120, 120)), ("c2c", (96, 96, 96)),
                   ("c2c", (96, 96, 96)), ("c2c", (32, 32, 32, 32))]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["grouped", "each"])
def test_batch_workload_shapes_vs_oracle(prec, mode):
    """The uneven workload's R2C and cube shapes, several per group."""
    from parafaft_amd.batch import Batch
    cdt = np.complex64 if prec == "f32" else np.complex128
    rdt = np.float32 if prec == "f32" else np.float64
    items = WORKLOAD_SHAPES
    offs, total = _layout(items, cdt)
    xs = _inputs(items, 7)
    xs = [x.astype(cdt).astype(np.complex128) if k == "c2c" else x.astype(rdt).astype(np.float64)
          for (k, _), x in zip(items, xs)]
    host = _pack_inputs(items, xs, offs, total, cdt, rdt)
    refs = _refs(items, xs)
    d = to_dev(host)
    esz = np.dtype(cdt).itemsize
    b = Batch(prec, [(k, s, d.data_ptr() + o * esz) for (k, s), o in zip(items, offs)])
    (b.execute if mode == "grouped" else b.execute_each)(0)
    _check(items, to_host(d), offs, refs, prec, cdt)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batch_tuned_vs_oracle(prec):
    """pfx_batch_tune (Measure) on the transforms' own buffers, then the
    regrouped schedule forward and backward against the oracle; grouped and
    one-launch-per-stage execution of the tuned plans agree bit for bit
    (same kernels, expression-level contraction)."""
    from parafaft_amd import _capi as A
    from parafaft_amd.batch import Batch
    cdt = np.complex64 if prec == "f32" else np.complex128
    rdt = np.float32 if prec == "f32" else np.float64
    esz = np.dtype(cdt).itemsize
    offs, total = _layout(SMALL, cdt)
    xs = _inputs(SMALL, 140)
    host = _pack_inputs(SMALL, xs, offs, total, cdt, rdt)
    refs = _refs(SMALL, [x.astype(cdt if k == "c2c" else rdt).astype(
        np.complex128 if k == "c2c" else np.float64) for (k, _), x in zip(SMALL, xs)])
    d = to_dev(host)
    b = Batch(prec, [(k, s, d.data_ptr() + o * esz) for (k, s), o in zip(SMALL, offs)])
    b.tune(A.PFX_PLAN_MEASURE)
    outs = {}
    for mode in ("grouped", "each"):
        d.copy_(torch().from_numpy(host).to("cuda"))
        torch().cuda.synchronize()
        (b.execute if mode == "grouped" else b.execute_each)(0)
        outs[mode] = to_host(d)
        _check(SMALL, outs[mode], offs, refs, prec, cdt)
    assert np.array_equal(outs["grouped"].view(np.uint8), outs["each"].view(np.uint8))


def test_batch_chunked_schedule(monkeypatch):
    """Cache-sized chunks (forced tiny here: every transform of 64 KiB or
    more is chunked, 1 MiB per chunk): the chunked grouped schedule gives
    the same bytes as one launch per stage, and matches the oracle."""
    from parafaft_amd.batch import Batch
    monkeypatch.setenv("PFX_BATCH_SMALL_MB", "0")
    monkeypatch.setenv("PFX_BATCH_CHUNK_MB", "1")
    prec, cdt, rdt = "f32", np.complex64, np.float32
    offs, total = _layout(SMALL, cdt)
    xs = _inputs(SMALL, 230)
    host = _pack_inputs(SMALL, xs, offs, total, cdt, rdt)
    refs = _refs(SMALL, [x.astype(cdt if k == "c2c" else rdt).astype(
        np.complex128 if k == "c2c" else np.float64) for (k, _), x in zip(SMALL, xs)])
    d = to_dev(host)
    b = Batch(prec, [(k, s, d.data_ptr() + o * 8) for (k, s), o in zip(SMALL, offs)])
    g, e = b.launches(0)
    monkeypatch.delenv("PFX_BATCH_SMALL_MB")
    monkeypatch.delenv("PFX_BATCH_CHUNK_MB")
    b_one = Batch(prec, [(k, s, d.data_ptr() + o * 8) for (k, s), o in zip(SMALL, offs)])
    g_one, _ = b_one.launches(0)
    assert g > g_one   # chunking splits the levels
    outs = {}
    for mode in ("chunked", "each"):
        d.copy_(torch().from_numpy(host).to("cuda"))
        torch().cuda.synchronize()
        (b.execute if mode == "chunked" else b.execute_each)(0)
        outs[mode] = to_host(d)
        _check(SMALL, outs[mode], offs, refs, prec, cdt)
    assert np.array_equal(outs["chunked"].view(np.uint8), outs["each"].view(np.uint8))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_batch_streaming_grouped(monkeypatch, prec):
    """The streaming (non-temporal) batch-contiguous kernels in their grouped
    form: every column pass with a streaming variant is forced onto it
    (PFX_KERNEL_MATCH), and the grouped schedule matches one launch per stage
    bit for bit and the oracle."""
    import parafaft_amd as P
    from parafaft_amd import backend as B
    from parafaft_amd.batch import Batch
    items = ([("c2c", (512, 64)), ("c2c", (512, 8, 8)), ("c2c", (1024, 36)), ("c2c", (512, 40)),
              ("c2c", (96, 512))] if prec == "f32" else
             [("c2c", (1024, 20)), ("c2c", (512, 24)), ("c2c", (1024, 8, 4))])
    cdt = np.complex64 if prec == "f32" else np.complex128
    rdt = np.float32 if prec == "f32" else np.float64
    esz = np.dtype(cdt).itemsize
    B.tune_cache_clear()   # a geometry tuned by an earlier test keeps its choice
    monkeypatch.setenv("PFX_KERNEL_MATCH", "wide-nt:")
    for kind, shape in items[:3]:
        plan = P.Plan.c2c(prec, shape[0], int(np.prod(shape[1:])), int(np.prod(shape[1:])), 1)
        assert "wide-nt" in plan.describe()[0], (shape, plan.describe())
    offs, total = _layout(items, cdt)
    xs = _inputs(items, 310)
    host = _pack_inputs(items, xs, offs, total, cdt, rdt)
    refs = _refs(items, [x.astype(cdt).astype(np.complex128) for x in xs])
    d = to_dev(host)
    b = Batch(prec, [(k, s, d.data_ptr() + o * esz) for (k, s), o in zip(items, offs)])
    monkeypatch.delenv("PFX_KERNEL_MATCH")
    outs = {}
    for mode in ("grouped", "each"):
        d.copy_(torch().from_numpy(host).to("cuda"))
        torch().cuda.synchronize()
        (b.execute if mode == "grouped" else b.execute_each)(0)
        outs[mode] = to_host(d)
        _check(items, outs[mode], offs, refs, prec, cdt)
    assert np.array_equal(outs["grouped"].view(np.uint8), outs["each"].view(np.uint8))


# radix-16 (and FP16 radix-32) transforms among the split-precision items
SPLIT_ITEMS = [("c2c", (256, 256)), ("c2c", (1024, 64)), ("c2c", (64, 1024)),
               ("c2c", (32, 32, 256)), ("r2c", (20, 24, 16)), ("c2c", (48, 48, 48)),
               ("c2c", (2, 1024, 40))]


@pytest.mark.parametrize("prec", ["f32_split_f16", "f32_split_bf16"])
@pytest.mark.parametrize("tuned", [False, True])
def test_batch_split_precision_vs_oracle(prec, tuned):
    """A batch in a split precision: its matrix-core stages have no grouped
    form and run as direct launches beside the grouped FP32 ones (lengths
    without radix-16 passes and R2C stages use the plain kernels); grouped and
    one launch per stage, forward and backward, every transform against the
    oracle at the FP32 bound."""
    from parafaft_amd.batch import Batch
    cdt, rdt = np.complex64, np.float32
    items = SPLIT_ITEMS
    offs, total = _layout(items, cdt)
    xs = _inputs(items, 70)
    host = _pack_inputs(items, xs, offs, total, cdt, rdt)
    refs = _refs(items, [x.astype(cdt if k == "c2c" else rdt).astype(
        np.complex128 if k == "c2c" else np.float64) for (k, _), x in zip(items, xs)])
    for mode in ("grouped", "each"):
        d = to_dev(host)
        b = Batch(prec, [(k, s, d.data_ptr() + o * 8) for (k, s), o in zip(items, offs)])
        if tuned:
            b.tune()
            d.copy_(to_dev(host))
        (b.execute if mode == "grouped" else b.execute_each)(0)
        _check(items, to_host(d), offs, refs, "f32", cdt)
        (b.execute if mode == "grouped" else b.execute_each)(1)
        back = to_host(d)
        for (kind, shape), o, x in zip(items, offs, xs):
            N = int(np.prod(shape))
            if kind == "c2c":
                got = back[o:o + x.size].reshape(shape)
            else:
                n = shape[-1]
                h = n // 2 + 1
                got = back[o:o + int(np.prod(shape[:-1])) * h].view(rdt).reshape(
                    shape[:-1] + (2 * h,))[..., :n]
            assert rel_l2(got, N * x) <= 1e-6 * 25, (prec, mode, kind, shape)
