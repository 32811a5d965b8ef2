"""One rank's ParaFaFT_R2C pencil share on one GPU (parafaft_r2c.hpp:360-430 with
exchange_packed's pack/unpack, parafaft_common.hpp:564-669), peers' blocks
left on the device as seeded synthetic segments.

- pack and unpack are checked bit-exact against the oracle's restatement of
  the reference's loops applied to the GPU's own intermediate data;
- the whole share transform is checked against the oracle pipeline
  (FP64 restated R2C / pack / unpack / C2C) on scaled-down grids, and against
  the FP64 GPU path at the full 2048^3-on-4x4 share (512 x 512 x 2048); the
  input is k^-3-filtered noise at every size;
- the fused form (pack and unpack inside the transforms) is checked
  bit-exact against the stepwise path: final output, every peer's send
  segment and the rank's own receive segment, on 2D, 3D and 4D shares and
  at the full share.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import rel_l2, to_dev, to_host, torch

pytestmark = pytest.mark.gpu


def filtered_noise(shape, seed):
    """iid N(0,1) white noise filtered to P(k) ~ k^-3 on the share treated as
    a periodic box (numpy, FP64; data generation only)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape)
    X = np.fft.rfftn(x)
    ks = np.meshgrid(*[np.fft.fftfreq(n) * n for n in shape[:-1]],
                     np.fft.rfftfreq(shape[-1]) * shape[-1], indexing="ij")
    k = np.sqrt(sum(kk ** 2 for kk in ks))
    amp = np.zeros_like(k)
    amp[k > 0] = k[k > 0] ** -1.5
    return np.fft.irfftn(X * amp, s=shape, axes=list(range(len(shape))))


def filtered_noise_device(shape, seed):
    """filtered_noise at full share size, generated on the GPU (FP32): iid
    N(0,1) from torch's generator, the library's own R2C / C2R of the share
    as a periodic box with the k^-3/2 amplitude in between (data generation
    only; returned as a host array)."""
    import parafaft_amd as P
    t = torch()
    n0, n1, n2 = shape
    h = n2 // 2 + 1
    fft = P.ParaFaFT_R2C(list(shape), precision="f32", fast=True)
    buf = t.zeros(fft.get_required_output_size(), dtype=t.float32, device="cuda")
    g = t.Generator(device="cuda")
    g.manual_seed(seed)
    buf[:n0 * n1 * 2 * h].view(n0 * n1, 2 * h)[:, :n2] = \
        t.randn((n0 * n1, n2), generator=g, device="cuda")
    fft.forward_in_place(buf)
    spec = t.view_as_complex(buf[:n0 * n1 * 2 * h].view(n0, n1, h, 2))
    k0 = t.fft.fftfreq(n0, device="cuda", dtype=t.float64) * n0
    k1 = t.fft.fftfreq(n1, device="cuda", dtype=t.float64) * n1
    k2 = t.arange(h, device="cuda", dtype=t.float64)
    for i in range(n0):   # one plane at a time keeps the FP64 temporaries small
        k = t.sqrt(k0[i] ** 2 + k1[:, None] ** 2 + k2[None, :] ** 2)
        amp = t.where(k > 0, k.clamp_min(1.0) ** -1.5, t.zeros_like(k)).to(t.float32)
        spec[i] *= amp
    fft.backward_in_place(buf, scale=1.0 / (n0 * n1 * n2))
    out = buf[:n0 * n1 * 2 * h].view(n0 * n1, 2 * h)[:, :n2].cpu().numpy().reshape(shape)
    del buf, spec
    t.cuda.empty_cache()
    return out


def _fill_peers(share, seed, cdt):
    """Seeded synthetic blocks in every peer receive segment; returns them."""
    peers = {}
    for t, g in enumerate(share.geoms):
        for p in range(len(g.recv_counts)):
            if p == share.me[t]:
                continue
            vals = O.uniform_complex((g.recv_counts[p],), seed=seed + 100 * t + p).astype(cdt)
            addr, count = share.recv_segment(t, p)
            dev = to_dev(vals)
            from parafaft_amd import _capi as A
            A.check(A.load().pfx_memcpy_async(addr, dev.data_ptr(), count * vals.itemsize, None),
                    "copy")
            torch().cuda.synchronize()
            peers[(t, p)] = vals
    return peers


def _oracle_share(real, share, peers):
    """FP64 oracle pipeline for the share, with the same peer blocks."""
    D = share.D
    ins, outs = share.stage_shapes, share.stage_output_shapes
    n = real.shape[-1]
    h = n // 2 + 1
    pad = np.zeros(real.shape[:-1] + (2 * h,))
    pad[..., :n] = real
    O.r2c_inplace(pad.reshape(-1, 2 * h), n, int(np.prod(real.shape[:-1])), 1, 2 * h)
    src = pad.view(np.complex128).reshape(outs[0])
    for t in range(D - 1):
        g = share.geoms[t]
        send = O.pack(src, g, 16).view(np.complex128)
        recv = np.zeros(sum(g.recv_counts), np.complex128)
        for p in range(len(g.recv_counts)):
            seg = slice(g.recv_displs[p], g.recv_displs[p] + g.recv_counts[p])
            if p == share.me[t]:
                recv[seg] = send[g.send_displs[p]:g.send_displs[p] + g.send_counts[p]]
            else:
                recv[seg] = peers[(t, p)].astype(np.complex128)
        dst = np.zeros(ins[t + 1], np.complex128)
        O.unpack(recv.view(np.uint8), dst, g, 16)
        axis = D - 2 - t
        length = dst.shape[axis]
        trail = int(np.prod(dst.shape[axis + 1:]))
        lead = int(np.prod(dst.shape[:axis]))
        flat = dst.reshape(lead, length * trail)
        for i in range(lead):
            row = np.ascontiguousarray(flat[i])
            O.c2c_batched(row, length, trail, trail, 1, -1)
            flat[i] = row
        src = flat.reshape(ins[t + 1])
    return src


def _run_stepwise(share, real, rdt, check_bytes=True):
    """Runs the share forward step by step on the GPU, checking every pack and
    unpack bit-exact against the oracle loops on the GPU's own data."""
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    D = share.D
    n = real.shape[-1]
    h = n // 2 + 1
    esz = share.esz
    cdt = np.complex64 if esz == 8 else np.complex128
    pad = np.zeros(real.shape[:-1] + (2 * h,), rdt)
    pad[..., :n] = real
    buf = np.zeros(share.get_required_output_size(), rdt)
    buf[:pad.size] = pad.ravel()
    d = to_dev(buf)
    share.r2c.execute(A.PFX_FORWARD, d, None)
    src, dst = d.data_ptr(), share.scratch.ptr
    for t in range(D - 1):
        g = share.geoms[t]
        share.pack(t, src)
        if check_bytes:
            src_host = _download(src, math.prod(share.stage_output_shapes[t]) * esz)
            ref = O.pack(src_host, g, esz)
            got = _download(share.send[t].ptr, sum(g.send_counts) * esz)
            assert np.array_equal(got, ref), f"pack transition {t} not bit-exact"
        share.self_copy(t)
        share.unpack(t, dst)
        if check_bytes:
            recv_host = _download(share.recv[t].ptr, sum(g.recv_counts) * esz)
            ref = O.unpack(recv_host, np.zeros(math.prod(share.stage_shapes[t + 1]) * esz,
                                               np.uint8), g, esz)
            got = _download(dst, math.prod(share.stage_shapes[t + 1]) * esz)
            assert np.array_equal(got, ref), f"unpack transition {t} not bit-exact"
        share.fft_stage(t + 1, dst)
        src, dst = dst, src
    out = _download(src, math.prod(share.stage_shapes[D - 1]) * esz)
    return out.view(cdt).reshape(share.stage_shapes[D - 1])


def _download(addr, nbytes):
    t = torch()
    t.cuda.synchronize()
    host = t.empty(nbytes, dtype=t.uint8)
    from parafaft_amd import _capi as A
    A.check(A.load().pfx_memcpy_async(host.data_ptr(), addr, nbytes, None), "d2h")
    A.check(A.load().pfx_stream_synchronize(None), "sync")
    return host.numpy()


@pytest.mark.parametrize("coords", [(0, 0), (3, 1), (1, 2)])
@pytest.mark.parametrize("fast", [True, False])
def test_pencil_share_small_vs_oracle(coords, fast):
    """128^3 real grid on a 4x4 grid: share 32 x 32 x 128 (complex last
    axis 65 split 17/16/16/16, like 1025 in the 2048^3 case)."""
    import parafaft_amd as P
    from parafaft_amd.pencil import PencilShareR2C
    G = [128, 128, 128]
    for prec, rdt, cdt, tol in (("f32", np.float32, np.complex64, 1e-6 * 21),
                                ("f64", np.float64, np.complex128, 1e-13)):
        share = PencilShareR2C(G, [4, 4], list(coords), precision=prec, fast=fast,
                               fused=False)
        real = filtered_noise(tuple(share.stage_shapes[0]), seed=31).astype(rdt)
        peers = _fill_peers(share, 500, cdt)
        got = _run_stepwise(share, real, rdt)
        ref = _oracle_share(real.astype(np.float64), share, peers)
        assert rel_l2(got, ref) <= tol, (prec, coords, rel_l2(got, ref))


def test_geometry_matches_oracle():
    from parafaft_amd import geometry as G
    for gshape, dims, coords in [([2048] * 3, [4, 4], [0, 0]), ([2048] * 3, [4, 4], [3, 1]),
                                 ([64, 48, 30], [2, 3], [1, 2]), ([16] * 4, [2, 2, 2], [1, 0, 1])]:
        assert list(G.r2c_stage_shapes(gshape, dims, coords)) == \
            list(O.stage_shapes_r2c(gshape, dims, coords))
    g1 = G.exchange_geometry(4, [512, 512, 1025], 2, [512, 2048, 257], 1)
    g2 = O.exchange_geometry(4, [512, 512, 1025], 2, [512, 2048, 257], 1)
    assert g1.__dict__ == g2.__dict__


def test_config3_full_share():
    """Rank (0,0) of a 2048^3 R2C on a 4x4 grid: real share 512 x 512 x 2048
    (padded 2050) of k^-3-filtered noise, half spectrum 512 x 512 x 1025; T0
    and T1 pack/unpack into 4 + 4 device segments each, checked bit-exact on
    the GPU's own data; the FP32 result against the FP64 GPU path on the
    same input (that path is pinned to the oracle above at <= 1e-13)."""
    from parafaft_amd.pencil import PencilShareR2C
    G = [2048, 2048, 2048]
    share32 = PencilShareR2C(G, [4, 4], [0, 0], precision="f32", fused=False)
    assert share32.stage_shapes == [[512, 512, 2048], [512, 2048, 257], [2048, 512, 257]]
    assert len(share32.geoms) == 2 and all(len(g.send_counts) == 4 for g in share32.geoms)
    real = filtered_noise_device((512, 512, 2048), 77)
    assert np.isfinite(real).all() and real.std() > 0
    _fill_peers(share32, 900, np.complex64)
    got32 = _run_stepwise(share32, real, np.float32, check_bytes=True)
    del share32
    share64 = PencilShareR2C(G, [4, 4], [0, 0], precision="f64", fused=False)
    _fill_peers(share64, 900, np.complex128)
    got64 = _run_stepwise(share64, real.astype(np.float64), np.float64, check_bytes=False)
    err = rel_l2(got32, got64)
    assert err <= 1e-6 * 11, err


def _segments(share):
    """Host bytes of every peer's send segment and of the rank's own receive
    segment, per transition (what leaves the rank and what it keeps)."""
    out = {}
    for t, g in enumerate(share.geoms):
        for p in range(len(g.send_counts)):
            if p == share.me[t]:
                addr, count = share.recv_segment(t, p)
                out[(t, "self")] = _download(addr, count * share.esz)
            else:
                addr, count = share.send_segment(t, p)
                out[(t, p)] = _download(addr, count * share.esz)
    return out


def _run_fused(share, real, rdt):
    n = real.shape[-1]
    h = n // 2 + 1
    pad = np.zeros(real.shape[:-1] + (2 * h,), rdt)
    pad[..., :n] = real
    buf = np.zeros(share.get_required_output_size(), rdt)
    buf[:pad.size] = pad.ravel()
    d = to_dev(buf)
    share.forward_in_place(d)
    nbytes = math.prod(share.stage_shapes[share.D - 1]) * share.esz
    return _download(d.data_ptr(), nbytes)


FUSED_CASES = [
    ([128, 128, 128], [4, 4], [3, 1]),
    ([128, 128, 128], [4, 4], [0, 2]),
    ([96, 80, 64], [2, 3], [1, 2]),
    ([64, 256], [4], [1]),
    ([32, 16, 32, 64], [2, 2, 2], [1, 0, 1]),
    # complex last axis 3 over 4 ranks: blocks 1, 1, 1, 0 (an empty peer block)
    ([8, 8, 4], [4, 4], [1, 2]),
    # ... and this rank's own complex share is empty
    ([8, 8, 4], [4, 4], [1, 3]),
]


@pytest.mark.parametrize("case", range(len(FUSED_CASES)))
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_fused_exchange_bit_exact(case, prec):
    from parafaft_amd.pencil import PencilShareR2C
    G, dims, coords = FUSED_CASES[case]
    rdt, cdt = (np.float32, np.complex64) if prec == "f32" else (np.float64, np.complex128)
    ref_share = PencilShareR2C(G, dims, coords, precision=prec, fused=False)
    fused = PencilShareR2C(G, dims, coords, precision=prec)
    assert fused.fused and fused.scratch is None
    real = filtered_noise(tuple(ref_share.stage_shapes[0]), seed=41 + case).astype(rdt)
    peers = _fill_peers(ref_share, 700, cdt)
    _fill_peers(fused, 700, cdt)
    want = _run_stepwise(ref_share, real, rdt, check_bytes=False)
    got = _run_fused(fused, real, rdt)
    a, b = _segments(ref_share), _segments(fused)
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), f"segment {k} differs"
    assert np.array_equal(got, want.view(np.uint8).ravel()), "fused output differs"
    if prec == "f64":
        ref = _oracle_share(real.astype(np.float64), ref_share, peers)
        assert rel_l2(got.view(cdt).reshape(ref.shape), ref) <= 1e-13


def test_fused_full_share_bit_exact():
    """The 512 x 512 x 2048 share of 2048^3 on 4x4, k^-3-filtered noise,
    FP32: fused vs stepwise."""
    from parafaft_amd.pencil import PencilShareR2C
    G = [2048, 2048, 2048]
    real = filtered_noise_device((512, 512, 2048), 78)
    ref_share = PencilShareR2C(G, [4, 4], [1, 2], precision="f32", fused=False)
    _fill_peers(ref_share, 910, np.complex64)
    want = _run_stepwise(ref_share, real, np.float32, check_bytes=False).view(np.uint8).ravel()
    segs_want = _segments(ref_share)
    del ref_share
    fused = PencilShareR2C(G, [4, 4], [1, 2], precision="f32")
    _fill_peers(fused, 910, np.complex64)
    got = _run_fused(fused, real, np.float32)
    assert np.array_equal(got, want)
    segs = _segments(fused)
    for k in segs_want:
        assert np.array_equal(segs[k], segs_want[k]), f"segment {k} differs"


def test_fused_rejects_bad_maps():
    """Host-side checks of pfx_execute_segmented: blocks must tile the axis."""
    import ctypes as C
    from parafaft_amd import _capi as A
    from parafaft_amd.backend import Plan
    plan = Plan.c2c(A.PFX_F32, 64, 4, 1, 64)
    d = torch().zeros(64 * 4 * 2, dtype=torch().float32, device="cuda")
    segs = (C.c_void_p * 2)(d.data_ptr(), d.data_ptr())
    m = A.SegMap(2, segs, (C.c_int * 2)(0, 30), (C.c_int * 2)(30, 30), 1)
    rc = A.load().pfx_execute_segmented(plan.handle, A.PFX_FORWARD, d.data_ptr(), C.byref(m),
                                        None, 1.0, None)
    assert rc == A.PFX_ERR_INVALID_ARG
    m = A.SegMap(2, segs, (C.c_int * 2)(0, 33), (C.c_int * 2)(32, 32), 1)
    rc = A.load().pfx_execute_segmented(plan.handle, A.PFX_FORWARD, d.data_ptr(), C.byref(m),
                                        None, 1.0, None)
    assert rc == A.PFX_ERR_INVALID_ARG


@pytest.mark.parametrize("fused", [True, False])
def test_share_tuned_vs_oracle(fused):
    """Measure planning of the share (every candidate kernel timed inside the
    whole forward), then the forward against the FP64 oracle pipeline."""
    from parafaft_amd.pencil import PencilShareR2C
    G = [128, 128, 128]
    share = PencilShareR2C(G, [4, 4], [1, 2], precision="f32", fused=fused)
    scratch = to_dev(np.zeros(share.get_required_output_size(), np.float32))
    kernels = share.tune(scratch)
    assert len(kernels) == 3
    real = filtered_noise(tuple(share.stage_shapes[0]), seed=77).astype(np.float32)
    peers = _fill_peers(share, 640, np.complex64)
    got = _run_fused(share, real, np.float32) if fused else \
        _run_stepwise(share, real, np.float32, check_bytes=False).view(np.uint8).ravel()
    ref = _oracle_share(real.astype(np.float64), share, peers)
    assert rel_l2(got.view(np.complex64).reshape(ref.shape), ref) <= 1e-6 * 21
