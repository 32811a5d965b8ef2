"""The seeded uneven workload at full size: 512 independent transforms, 40%
C2C cubes {48,64,96,128}^3, 30% C2C squares {1024,2048,4096}^2, 20% R2C cubes
{80,120,160}^3, 10% C2C 32^4, complex<float>, inputs iid Uniform[-1,1).

Every schedule of the batch runs on the device: grouped launches, the
grouped schedule replayed as a HIP graph, one launch per transform stage
and the persistent single launch.  Every transform is checked against the
FP64 path on the same input (rel-L2 <= 1e-6 log2 N, computed on the GPU),
and one transform of every shape class, FP32 and FP64, against the oracle
(the pattern of the reference's test/c2c/test_fftw_c2c_nd.cpp:98-119, which
compares every ParaFaFT output with a serial FFTW plan).  The oracle's
reference is computed per transform and dropped before the next.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, torch

pytestmark = pytest.mark.gpu

FORWARD, BACKWARD = 0, 1


def _layout(items):
    from parafaft_amd.batch import storage_elems
    offs, pos = [], 0
    for kind, shape in items:
        offs.append(pos)
        pos += (storage_elems(kind, shape) + 63) // 64 * 64
    return offs, pos


@pytest.fixture(scope="module")
def wl():
    """The workload, its device input and the FP64 path's forward result."""
    from parafaft_amd.batch import Batch, storage_elems, uneven_workload
    t = torch()
    items = uneven_workload()
    assert len(items) == 512
    offs, total = _layout(items)
    g = t.Generator(device="cuda")
    g.manual_seed(11)
    x0 = t.view_as_complex(t.rand((total, 2), generator=g, device="cuda") * 2 - 1)
    y64 = x0.to(t.complex128)
    b64 = Batch("f64", [(k, s, y64.data_ptr() + o * 16) for (k, s), o in zip(items, offs)])
    b64.execute(FORWARD)
    t.cuda.synchronize()
    del b64
    sizes = [storage_elems(k, s) for k, s in items]
    yield {"items": items, "offs": offs, "sizes": sizes, "x0": x0, "y64": y64}
    del x0, y64
    t.cuda.empty_cache()


def _batch(wl, buf):
    from parafaft_amd.batch import Batch
    return Batch("f32", [(k, s, buf.data_ptr() + o * 8)
                         for (k, s), o in zip(wl["items"], wl["offs"])])


def _check_vs_fp64(wl, y):
    """Every transform's FP32 result against the FP64 path (on the GPU)."""
    t = torch()
    worst = 0.0
    for (kind, shape), o, n in zip(wl["items"], wl["offs"], wl["sizes"]):
        ref = wl["y64"][o:o + n]
        e = float(t.linalg.vector_norm(y[o:o + n].to(t.complex128) - ref) /
                  t.linalg.vector_norm(ref))
        assert e <= 1e-6 * math.log2(math.prod(shape)), (kind, shape, o, e)
        worst = max(worst, e)
    return worst


def _classes(items):
    """Index of the first transform of every (kind, shape) class."""
    first = {}
    for i, (kind, shape) in enumerate(items):
        first.setdefault((kind, tuple(shape)), i)
    return first


def _host_input(wl, i):
    """Transform i's input as the oracle takes it (FP32 values, in FP64)."""
    kind, shape = wl["items"][i]
    o, n = wl["offs"][i], wl["sizes"][i]
    x = wl["x0"][o:o + n].cpu().numpy()
    if kind == "c2c":
        return x.astype(np.complex128).reshape(shape)
    h = shape[-1] // 2 + 1
    real = x.view(np.float32).reshape(tuple(shape[:-1]) + (2 * h,))[..., :shape[-1]]
    return real.astype(np.float64)


def _oracle(kind, x):
    return O.fftn(x) if kind == "c2c" else O.rfftn_padded(x)


def test_every_shape_class_fp64_path_vs_oracle(wl):
    """The FP64 path (the per-transform reference of the other tests) against
    the oracle, one transform of each of the 11 shape classes."""
    classes = _classes(wl["items"])
    assert len(classes) == 11, sorted(classes)
    for (kind, shape), i in sorted(classes.items()):
        ref = _oracle(kind, _host_input(wl, i))
        o = wl["offs"][i]
        got = wl["y64"][o:o + ref.size].cpu().numpy().reshape(ref.shape)
        assert rel_l2(got, ref) <= 1e-13, (kind, shape, rel_l2(got, ref))
        del ref, got


def test_grouped_forward_every_transform(wl):
    """Grouped launches: every transform against the FP64 path, one of every
    class against the oracle; the backward of the result returns N x."""
    t = torch()
    x = wl["x0"].clone()
    b = _batch(wl, x)
    grouped, each = b.launches(FORWARD)
    assert grouped < each
    b.execute(FORWARD)
    t.cuda.synchronize()
    worst = _check_vs_fp64(wl, x)
    print(f"grouped: {grouped} launches ({each} one per stage), worst rel-L2 vs FP64 {worst:.2e}",
          flush=True)
    for (kind, shape), i in sorted(_classes(wl["items"]).items()):
        ref = _oracle(kind, _host_input(wl, i))
        o = wl["offs"][i]
        got = x[o:o + ref.size].cpu().numpy().reshape(ref.shape)
        e = rel_l2(got, ref)
        assert e <= 1e-6 * math.log2(math.prod(shape)), (kind, shape, e)
        del ref, got
    # backward: C2C gives N x; R2C gives N x in the real slots of each row
    b.execute(BACKWARD)
    t.cuda.synchronize()
    for (kind, shape), o, n in zip(wl["items"], wl["offs"], wl["sizes"]):
        N = math.prod(shape)
        if kind == "c2c":
            got, ref = x[o:o + n], wl["x0"][o:o + n]
        else:
            h = shape[-1] // 2 + 1
            rows = n // h
            got = t.view_as_real(x[o:o + n]).reshape(rows, 2 * h)[:, :shape[-1]]
            ref = t.view_as_real(wl["x0"][o:o + n]).reshape(rows, 2 * h)[:, :shape[-1]]
        got64 = got.to(t.complex128 if kind == "c2c" else t.float64)
        ref64 = ref.to(got64.dtype) * N
        e = float(t.linalg.vector_norm(got64 - ref64) / t.linalg.vector_norm(ref64))
        assert e <= 2e-6 * math.log2(N), (kind, shape, e)
    del x, b


def test_each_graph_bit_identical_to_grouped(wl):
    """One launch per transform stage, and the grouped schedule replayed as a
    HIP graph (three replays with eager work between), give the grouped
    launches' bytes: the same kernels on the same data."""
    from parafaft_amd.batch import Graph
    t = torch()
    x = wl["x0"].clone()
    b = _batch(wl, x)
    b.execute(FORWARD)
    t.cuda.synchronize()
    grouped = x.clone()
    x.copy_(wl["x0"])
    t.cuda.synchronize()
    b.execute_each(FORWARD)
    t.cuda.synchronize()
    assert t.equal(t.view_as_real(x), t.view_as_real(grouped))
    stream = t.cuda.Stream()
    sp = stream.cuda_stream
    graph = Graph(sp, lambda: b.execute(FORWARD, sp))
    for rep in range(3):
        x.copy_(wl["x0"])
        t.cuda.synchronize()
        _ = (t.ones(1000, device="cuda") * 2).sum().item()
        graph.launch(sp)
        stream.synchronize()
        assert t.equal(t.view_as_real(x), t.view_as_real(grouped)), rep
    del graph, x, b, grouped


def test_persistent_single_launch(wl):
    """The whole workload in one persistent launch: bit-identical to the
    grouped launches (the persistent bodies use the default plans' radices),
    so every transform also meets the FP64 bound checked above; the backward
    launch likewise."""
    t = torch()
    x = wl["x0"].clone()
    b = _batch(wl, x)
    jobs, tickets, wgs = b.persistent_info(FORWARD)
    assert jobs == sum(len(s) for _, s in wl["items"])
    print(f"persistent: {jobs} jobs, {tickets} tickets, {wgs} workgroups", flush=True)
    b.execute(FORWARD)
    t.cuda.synchronize()
    grouped = x.clone()
    x.copy_(wl["x0"])
    t.cuda.synchronize()
    b.execute_persistent(FORWARD)
    assert not b.persistent_failed()
    assert t.equal(t.view_as_real(x), t.view_as_real(grouped))
    _check_vs_fp64(wl, x)
    b.execute(BACKWARD)
    t.cuda.synchronize()
    grouped_back = x.clone()
    x.copy_(grouped)
    t.cuda.synchronize()
    b.execute_persistent(BACKWARD)
    assert not b.persistent_failed()
    assert t.equal(t.view_as_real(x), t.view_as_real(grouped_back))
    del x, b, grouped, grouped_back
