"""The persistent single launch of a batch of uneven transforms
(pfx_batch_execute_persistent): tickets from per-XCD queues, stage
dependencies counted on the device, every wait bounded.

Checked against the grouped schedule (bit-identical: the persistent bodies
use the default plans' radices) and against the oracle, forward and
backward, repeated launches and graph replay; the refusals (lengths without
a radix plan, complex<double>); and the bounded wait, forced from the host
with a tiny spin limit (every workgroup still exits normally)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, torch
from tests.test_gpu_batch import _inputs, _layout, _pack_inputs, _refs

pytestmark = pytest.mark.gpu

# every body family: LDS rows and columns of small and mixed-radix lengths,
# register-resident rows (4096, 2048, 1024) and columns (1024, 2048, 4096),
# R2C/C2R of even real lengths, 1D to 4D
SHAPES = [("c2c", (48, 48, 48)), ("c2c", (64, 64, 64)), ("c2c", (16, 24)), ("c2c", (1024, 32)),
          ("r2c", (20, 24, 16)), ("c2c", (8, 8, 8, 8)), ("c2c", (96,)), ("r2c", (80, 12, 10)),
          ("c2c", (120, 8, 10)), ("c2c", (32, 32, 32, 32)), ("c2c", (256, 256)),
          ("c2c", (64, 4096)), ("c2c", (2048, 40)), ("r2c", (160, 8, 6)), ("c2c", (4096, 64)),
          ("c2c", (40, 2048)), ("c2c", (24, 1024)), ("r2c", (6, 4, 4096)), ("c2c", (5, 3, 60))]


def _setup(items, seed, prec="f32"):
    cdt = np.complex64 if prec == "f32" else np.complex128
    rdt = np.float32 if prec == "f32" else np.float64
    offs, total = _layout(items, cdt)
    xs = _inputs(items, seed)
    xs = [x.astype(cdt).astype(np.complex128) if k == "c2c" else x.astype(rdt).astype(np.float64)
          for (k, _), x in zip(items, xs)]
    host = _pack_inputs(items, xs, offs, total, cdt, rdt)
    return offs, xs, host


def _batch(items, d, offs, prec="f32"):
    from parafaft_amd.batch import Batch
    esz = 8 if prec == "f32" else 16
    return Batch(prec, [(k, s, d.data_ptr() + o * esz) for (k, s), o in zip(items, offs)])


def test_persistent_forward_backward_vs_grouped_and_oracle():
    t = torch()
    items = SHAPES
    offs, xs, host = _setup(items, 501)
    refs = _refs(items, xs)
    d = to_dev(host)
    b = _batch(items, d, offs)
    jobs, tickets, wgs = b.persistent_info(0)
    assert jobs == sum(len(s) for _, s in items)
    assert tickets >= jobs and wgs >= 8
    print(f"persistent: {jobs} jobs, {tickets} tickets, {wgs} workgroups", flush=True)
    out = {}
    for mode in ("grouped", "persistent"):
        d.copy_(t.from_numpy(host).to("cuda"))
        t.cuda.synchronize()
        (b.execute if mode == "grouped" else b.execute_persistent)(0)
        t.cuda.synchronize()
        if mode == "persistent":
            assert not b.persistent_failed()
        out[mode] = to_host(d)
    assert np.array_equal(out["grouped"].view(np.uint8), out["persistent"].view(np.uint8))
    for (kind, shape), o, r in zip(items, offs, refs):
        got = out["persistent"][o:o + r.size].reshape(r.shape)
        assert rel_l2(got, r) <= 1e-6 * 22, (kind, shape, rel_l2(got, r))
    # backward of the spectrum: N x (R2C: in the real slots), both schedules alike
    back = {}
    for mode in ("grouped", "persistent"):
        d.copy_(t.from_numpy(out["grouped"]).to("cuda"))
        t.cuda.synchronize()
        (b.execute if mode == "grouped" else b.execute_persistent)(1)
        t.cuda.synchronize()
        if mode == "persistent":
            assert not b.persistent_failed()
        back[mode] = to_host(d)
    assert np.array_equal(back["grouped"].view(np.uint8), back["persistent"].view(np.uint8))
    for (kind, shape), o, x in zip(items, offs, xs):
        N = int(np.prod(shape))
        if kind == "c2c":
            got = back["persistent"][o:o + x.size].reshape(shape)
        else:
            n, h = shape[-1], shape[-1] // 2 + 1
            got = back["persistent"][o:o + int(np.prod(shape[:-1])) * h].view(np.float32).reshape(
                shape[:-1] + (2 * h,))[..., :n]
        assert rel_l2(got, N * x) <= 1e-6 * 25, (kind, shape)


def test_persistent_repeated_and_graph_replay():
    """The launch clears its own counters: repeated launches and replays of
    a captured graph (with eager work between) give the same bytes."""
    from parafaft_amd.batch import Graph
    t = torch()
    items = SHAPES[:12]
    offs, xs, host = _setup(items, 733)
    d = to_dev(host)
    b = _batch(items, d, offs)
    b.execute_persistent(0)
    first = to_host(d)
    for rep in range(3):
        d.copy_(t.from_numpy(host).to("cuda"))
        t.cuda.synchronize()
        b.execute_persistent(0)
        assert np.array_equal(to_host(d).view(np.uint8), first.view(np.uint8)), rep
    stream = t.cuda.Stream()
    sp = stream.cuda_stream
    g = Graph(sp, lambda: b.execute_persistent(0, sp))
    for rep in range(3):
        d.copy_(t.from_numpy(host).to("cuda"))
        t.cuda.synchronize()
        _ = (t.ones(1000, device="cuda") * 2).sum().item()
        g.launch(sp)
        stream.synchronize()
        assert np.array_equal(to_host(d).view(np.uint8), first.view(np.uint8)), rep
    assert not b.persistent_failed(sp)


def test_persistent_refusals():
    """Stage lengths without a radix plan (the direct-DFT fallback) and
    complex<double> batches have no persistent schedule: refused with
    PFX_ERR_UNSUPPORTED before any launch; the grouped schedule still runs."""
    import parafaft_amd as P
    items = [("c2c", (7, 11, 13)), ("c2c", (64, 64))]
    offs, xs, host = _setup(items, 3)
    d = to_dev(host)
    b = _batch(items, d, offs)
    with pytest.raises(P.PfxError) as e:
        b.execute_persistent(0)
    assert e.value.status == 2 and "without a persistent kernel body" in str(e.value)
    b.execute(0)
    refs = _refs(items, xs)
    got = to_host(d)
    for (kind, shape), o, r in zip(items, offs, refs):
        assert rel_l2(got[o:o + r.size].reshape(r.shape), r) <= 1e-6 * 14
    items64 = [("c2c", (64, 64))]
    offs64, _, host64 = _setup(items64, 4, "f64")
    d64 = to_dev(host64)
    b64 = _batch(items64, d64, offs64, "f64")
    with pytest.raises(P.PfxError) as e:
        b64.execute_persistent(0)
    assert e.value.status == 2


def test_persistent_bounded_wait_ends_the_launch():
    """A spin limit of zero polls, set from the host: a workgroup that finds
    a dependency unfinished gives up at once, raises the abort word, and
    every workgroup exits; the status call reports it.  The next launch with
    the default limit is complete and correct (the counters are cleared per
    launch).  One transform of 16 row tiles and 16 column tiles on one XCD:
    the XCD's workgroups claim all 32 tickets at once, so half of them start
    on column tiles while every row tile is still running."""
    t = torch()
    items = [("c2c", (64, 4096))]
    offs, xs, host = _setup(items, 9)
    d = to_dev(host)
    b = _batch(items, d, offs)
    jobs, tickets, wgs = b.persistent_info(0)
    assert (jobs, tickets) == (2, 32) and wgs >= 256
    b.set_spin_limit(0)
    b.execute_persistent(0)
    assert b.persistent_failed()
    b.set_spin_limit(1 << 21)
    d.copy_(t.from_numpy(host).to("cuda"))
    t.cuda.synchronize()
    b.execute_persistent(0)
    assert not b.persistent_failed()
    got = to_host(d)
    d.copy_(t.from_numpy(host).to("cuda"))
    t.cuda.synchronize()
    b.execute(0)
    assert np.array_equal(got.view(np.uint8), to_host(d).view(np.uint8))
    ref = O.fftn(xs[0])
    o = offs[0]
    assert rel_l2(got[o:o + ref.size].reshape(ref.shape), ref) <= 1e-6 * 22
