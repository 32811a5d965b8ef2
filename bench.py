#!/usr/bin/env python3
"""Benchmark: in-place 3D C2C forward FFT of a 1024^3 complex<float> grid on
one MI355X through libpfx.so (one-rank fast path: one launch per axis).

Prints ONE JSON line.  ``value`` is GFLOP/s = 5 N log2(N) / t per forward
transform (N = 1024^3), inputs resident in HBM.  Also reported: achieved HBM
TB/s on algorithmic bytes (2 * D * N * 8 B per transform), rel-L2 / max-abs of
the FP32 result against the FP64 path on the same input, the roofline of the
dominant stage kernel (HIP events on the launch stream), and a CPU baseline
(the FP64 oracle on a bounded sample of the same per-axis transforms).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--no-check]
                  [--no-cpu] [--no-split] [--precision f32|f32_split_f16|f32_split_bf16]

With FP32 precision the same line also carries ``split_vs_plain``: both
split-precision MFMA forms planned, checked and timed on the same input.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D FFT GFLOP/s + HBM TB/s, 1024^3 fp32, rel-L2≤2e-6 vs FP64; split vs plain"
HBM_PEAK_GBS = 8000.0
ACC_BOUND = 2e-6          # rel-L2 against the FP64 path, every FP32 form


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n: int, frac_log2: int = 3, reps: int = 5):
    """The CPU column, timed with the reference benchmark's protocol
    (benchmark/bench_c2c.cpp:61-76: one warm-up forward+backward, then timed
    forward+backward pairs; here `reps` of them, median and mean +- sd) for
    both precisions of FFTWBackend (batch-parallel serial 1D plans per axis,
    fft_backend_fftw.hpp:387-424, threads as at :768-808).  FFTW is not
    installed, so the plans are the oracle's restated ones: its FP64 FFT
    core, and the same core built in float.  The sample is a bounded part of
    the 1024^3 workload: for each of the three axes, 1/2^frac_log2 of that
    axis' n*n length-n transforms with that axis' stride pattern; the times
    are scaled by 2^frac_log2 to one whole transform."""
    import ctypes.util
    import statistics
    import numpy as np
    from oracle import oracle as O
    fftw = {lib: ctypes.util.find_library(lib) for lib in ("fftw3", "fftw3f")}
    m = (n * n) >> frac_log2          # transforms per axis in the sample
    lead = m // n                     # slices of the middle axis
    threads = O.max_threads()
    flops_full = 5.0 * n ** 3 * math.log2(n ** 3)
    scale = 1 << frac_log2
    res = {}
    t_start = time.perf_counter()
    for prec in ("f64", "f32"):
        cdt = np.complex128 if prec == "f64" else np.complex64
        run = O.c2c_batched if prec == "f64" else O.c2c_batched_f32
        buf = O.uniform_complex((m * n,), seed=12345).astype(cdt)

        def passes(sign):
            axes = [lambda: run(buf, n, m, 1, n, sign, threads),                # axis 2: rows
                    lambda: [run(buf[i * n * n:(i + 1) * n * n], n, n, n, 1, sign, threads)
                             for i in range(lead)],                                # axis 1
                    lambda: run(buf, n, m, m, 1, sign, threads)]                  # axis 0
            for f in (axes if sign < 0 else axes[::-1]):
                f()
        passes(-1)
        passes(+1)                    # warm-up forward+backward
        tf, tfb = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            passes(-1)
            t1 = time.perf_counter()
            passes(+1)
            t2 = time.perf_counter()
            tf.append((t1 - t0) * scale)
            tfb.append((t2 - t0) * scale)

        def summary(ts):
            return {"median_ms": round(statistics.median(ts) * 1e3, 2),
                    "mean_ms": round(statistics.mean(ts) * 1e3, 2),
                    "sd_ms": round(statistics.stdev(ts) * 1e3, 2)}
        res[prec] = {"forward": summary(tf), "forward_backward": summary(tfb),
                     "forward_gflops_median": round(flops_full / statistics.median(tf) / 1e9, 3)}
        del buf
    return {"value": res["f64"]["forward_gflops_median"], "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle's restated FFTWBackend<double> and <float> batched plans (OpenMP "
                       f"over the batch, {threads} threads): {m} of each axis' {n * n} length-{n} "
                       f"transforms with strides 1, {n} and {m} elements; one warm-up "
                       f"forward+backward, then {reps} timed pairs; times scaled x{scale} to one "
                       f"{n}^3 transform; {time.perf_counter() - t_start:.1f} s in all"),
            "fp64": res["f64"], "fp32": res["f32"],
            "fftw_libraries_found": fftw,
            "note": ("FFTW is absent on this machine (find_library found neither library), so "
                     "the column is the oracle's own plans, not FFTW" if not any(fftw.values())
                     else "FFTW libraries were found; the column still times the oracle's plans")}


def bench_pencil(args, torch, A):
    """Times ParaFaFT_R2C<3>::forward_in_place for rank (0,0) of a 2048^3 real
    grid on a 4x4 process grid: R2C on the 512x512x2048 share, then per
    transition pack -> self copy -> unpack -> C2C (--unfused), or the fused
    form where each stage stores into and loads from the exchange segments;
    peers' blocks resident."""
    from parafaft_amd.pencil import PencilShareR2C
    share = PencilShareR2C([2048] * 3, [4, 4], [0, 0], precision=args.precision,
                           fused=not args.unfused)
    rdt = torch.float64 if share.esz == 16 else torch.float32
    buf = torch.zeros(share.get_required_output_size(), dtype=rdt, device="cuda")
    kernels = None
    if args.plan != "estimate":
        t0 = time.perf_counter()
        kernels = share.tune(buf)     # Measure planning on the share's own buffer
        log(f"[bench] pencil planning: {time.perf_counter() - t0:.1f} s, kernels {kernels}")
        buf.zero_()
        torch.cuda.synchronize()
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    real = torch.randn((512 * 512, 2048), generator=g, device="cuda", dtype=rdt)
    buf[:512 * 512 * 2050].view(512 * 512, 2050)[:, :2048] = real
    del real
    # peers' incoming blocks: seeded synthetic values, resident on the device
    for t, geo in enumerate(share.geoms):
        for p in range(len(geo.recv_counts)):
            if p != share.me[t]:
                addr, cnt = share.recv_segment(t, p)
                blk = torch.randn(2 * cnt, generator=g, device="cuda", dtype=rdt)
                A.check(A.load().pfx_memcpy_async(addr, blk.data_ptr(), blk.numel() *
                                                  blk.element_size(), None), "peer fill")
                torch.cuda.synchronize()
                del blk
    torch.cuda.synchronize()
    if share.fused:
        labels = ["r2c", "fft1", "fft2", "end"]
    else:
        labels = ["r2c", "pack0", "unpack0", "fft1", "pack1", "unpack1", "fft2", "end"]
    stream = torch.cuda.current_stream()
    share.stream = stream.cuda_stream
    for _ in range(args.warmup):
        share.forward_in_place(buf)
    torch.cuda.synchronize()
    evs = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        marks = {}

        def mark(label, marks=marks):
            e = torch.cuda.Event(enable_timing=True)
            e.record(stream)
            marks[label] = e
        share.forward_in_place(buf, events=mark)
        evs.append(marks)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    parts = {}
    for a, b in zip(labels[:-1], labels[1:]):
        parts[a] = sum(m[a].elapsed_time(m[b]) for m in evs) / args.steps
    esz = share.esz
    real_bytes = 512 * 512 * 2048 * (esz // 2)
    cplx = 512 * 512 * 1025 * esz
    cplx2 = 512 * 2048 * 257 * esz
    alg = {"r2c": real_bytes + cplx, "pack0": 2 * cplx, "unpack0": 2 * cplx2, "fft1": 2 * cplx2,
           "pack1": 2 * cplx2, "unpack1": 2 * cplx2, "fft2": 2 * cplx2}
    alg = {k: v for k, v in alg.items() if k in parts}
    total = sum(alg.values())
    how = ("fused: stages store into / load from 4+4 device segments per transition"
           if share.fused else "pack/unpack into 4+4 device segments per transition")
    del buf
    torch.cuda.empty_cache()
    return ({
        "workload": "R2C pencil share: rank (0,0) of 2048^3 on a 4x4 grid, 512x512x2048 real, "
                    + how, "fused": share.fused,
        "precision": args.precision, "planning": args.plan, "stage_kernels": kernels,
        "steps": args.steps, "ms_per_forward": round(ms, 4),
        "part_ms": {k: round(v, 4) for k, v in parts.items()},
        "part_tbs": {k: round(alg[k] / (parts[k] * 1e-3) / 1e12, 3) for k in parts},
        "algorithmic_bytes": total, "tbs_algorithmic": round(total / (ms * 1e-3) / 1e12, 3),
        "gflops_3d": round(2.5 * 2048 ** 3 / 16 * math.log2(2048 ** 3) / (ms * 1e-3) / 1e9, 1)})


def bench_batch(args, torch, A):
    import parafaft_amd as P
    """512 uneven transforms (seeded draw: 40% {48..128}^3 C2C, 30% {1024..4096}^2
    C2C, 20% {80,120,160}^3 R2C, 10% 32^4 C2C), complex<float>: one launch per
    transform stage vs grouped launches vs the grouped schedule replayed as a
    HIP graph; per-transform rel-L2 against the FP64 path."""
    from parafaft_amd.batch import Batch, Graph, storage_elems, uneven_workload
    items = uneven_workload()
    offs, pos = [], 0
    for kind, shape in items:
        offs.append(pos)
        pos += (storage_elems(kind, shape) + 63) // 64 * 64
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    x0 = torch.view_as_complex(torch.rand((pos, 2), generator=g, device="cuda") * 2 - 1)
    x = x0.clone()
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    b = Batch("f32", [(k, s, x.data_ptr() + o * 8) for (k, s), o in zip(items, offs)])
    flag = {"estimate": A.PFX_PLAN_ESTIMATE, "measure": A.PFX_PLAN_MEASURE,
            "patient": A.PFX_PLAN_PATIENT}[args.plan]
    if flag != A.PFX_PLAN_ESTIMATE:
        t0 = time.perf_counter()
        b.tune(flag, sp)          # every distinct stage geometry timed once
        log(f"[bench] batch planning ({args.plan}): {time.perf_counter() - t0:.1f} s")
        torch.cuda.synchronize()
        x.copy_(x0)               # planning used the buffers as scratch
        torch.cuda.synchronize()
    grouped, each = b.launches(A.PFX_FORWARD)
    points = sum(math.prod(s) for _, s in items)
    flops = sum((5.0 if k == "c2c" else 2.5) * math.prod(s) * math.log2(math.prod(s))
                for k, s in items)
    res = {}

    def timed(label, fn, steps=None):
        steps = steps or args.steps
        for _ in range(args.warmup):
            fn()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        res[label] = {"ms": round(ms, 3), "gflops": round(flops / (ms * 1e-3) / 1e9, 1),
                      "gpoints_per_s": round(points / (ms * 1e-3) / 1e9, 3)}
        log(f"[bench] {label}: {ms:.3f} ms")

    timed("one_launch_per_transform_stage", lambda: b.execute_each(A.PFX_FORWARD, sp))
    timed("grouped", lambda: b.execute(A.PFX_FORWARD, sp))
    pinfo = None
    try:
        pinfo = b.persistent_info(A.PFX_FORWARD)
    except P.PfxError as e:
        log(f"[bench] no persistent schedule: {e}")
    if pinfo:
        timed("persistent", lambda: b.execute_persistent(A.PFX_FORWARD, sp))
        if b.persistent_failed(sp):
            raise RuntimeError("persistent launch stopped at its spin limit")
    graph = Graph(sp, lambda: b.execute(A.PFX_FORWARD, sp))
    timed("grouped_graph", lambda: graph.launch(sp))
    graph_each = Graph(sp, lambda: b.execute_each(A.PFX_FORWARD, sp))
    timed("each_graph", lambda: graph_each.launch(sp))
    del graph_each
    # The reference's own pattern: one ParaFaFT object per shape, driven as its
    # drivers drive a backend (per-slice middle-axis launches and the one-rank
    # exchange copies, parafaft_c2c.hpp:539-602, 866-897), on the same kernels;
    # replayed as one graph so host overhead is not counted.
    if not args.no_reference_driver:
        from parafaft_amd.parafaft import ParaFaFT_C2C, ParaFaFT_R2C
        from parafaft_amd.backend import FFTPlanFlag
        pf = FFTPlanFlag.Estimate if flag == A.PFX_PLAN_ESTIMATE else FFTPlanFlag.Measure
        drivers = []
        for (k, s), o in zip(items, offs):
            if k == "c2c":
                drivers.append((ParaFaFT_C2C(list(s), "f32", plan_flag=pf, fast=False,
                                             stream=sp), "c2c", x.data_ptr() + o * 8))
            else:
                drivers.append((ParaFaFT_R2C(list(s), "f32", plan_flag=pf, fast=False,
                                             stream=sp), "r2c", x.data_ptr() + o * 8))
        torch.cuda.synchronize()
        x.copy_(x0)
        torch.cuda.synchronize()

        def run_drivers():
            for drv, kind, addr in drivers:
                if kind == "c2c":
                    drv.forward(addr, sync=False)
                else:
                    drv.forward_in_place(addr, sync=False)
        t0 = time.perf_counter()
        graph_ref = Graph(sp, run_drivers)
        log(f"[bench] reference-pattern graph captured in {time.perf_counter() - t0:.1f} s")
        timed("reference_driver_per_shape_graph", lambda: graph_ref.launch(sp),
              steps=min(args.steps, 5))
        del graph_ref, drivers
    # accuracy: one grouped forward on fresh data vs the FP64 batch on the same input
    acc = None
    if not args.no_check:
        # torch's stream and `stream` do not order each other: synchronise
        torch.cuda.synchronize()
        x.copy_(x0)
        x64 = x0.to(torch.complex128)
        torch.cuda.synchronize()
        b.execute(A.PFX_FORWARD, sp)
        b64 = Batch("f64", [(k, s, x64.data_ptr() + o * 16) for (k, s), o in zip(items, offs)])
        b64.execute(A.PFX_FORWARD, sp)
        stream.synchronize()
        worst, worst_item = 0.0, None
        for (k, s), o in zip(items, offs):
            n = storage_elems(k, s)
            ref = x64[o:o + n]
            e = float(torch.linalg.vector_norm(x[o:o + n].to(torch.complex128) - ref) /
                      torch.linalg.vector_norm(ref))
            if e > worst:
                worst, worst_item = e, f"{k}{list(s)}"
        acc = {"max_rel_l2_vs_fp64": worst, "worst_transform": worst_item}
        del x64, b64
    del graph
    del x, x0, b
    torch.cuda.empty_cache()
    return ({
        "workload": "512 uneven transforms, complex<float>, seeded shapes", "transforms": len(items),
        "complex_points": points, "resident_gb": round(pos * 8 / 1e9, 3),
        "launches_per_forward": {"one_per_transform_stage": each, "grouped": grouped},
        "results": res,
        "speedup_grouped_vs_each": round(res["one_launch_per_transform_stage"]["ms"] /
                                         res["grouped"]["ms"], 3),
        "persistent": ({"jobs": pinfo[0], "tickets": pinfo[1], "workgroups": pinfo[2],
                        "speedup_vs_each": round(res["one_launch_per_transform_stage"]["ms"] /
                                                 res["persistent"]["ms"], 3)}
                       if pinfo else None),
        "speedup_graph_vs_each": round(res["one_launch_per_transform_stage"]["ms"] /
                                       res["grouped_graph"]["ms"], 3),
        "speedup_graph_vs_reference_driver": (
            round(res["reference_driver_per_shape_graph"]["ms"] / res["grouped_graph"]["ms"], 3)
            if "reference_driver_per_shape_graph" in res else None),
        "planning": args.plan,
        "accuracy": acc})


def bench_512(args, torch, A):
    """The 512^3 in-place C2C of the GPU plain-method case, complex<double> and
    complex<float>: iid Uniform[-1,1) re/im plus 8 plane waves
    exp(+2 pi i k_m . x / 512) with seeded integer wavevectors; Measure
    planning; K forwards timed.  Accuracy without the oracle (which only the
    tests load): the FP64 path's scaled roundtrip and Parseval's identity,
    and the FP32 path against the FP64 path on the same input (the tests pin
    the FP64 path to the oracle at this size)."""
    import ctypes as C
    n, N = 512, 512 ** 3
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(512)
    waves = torch.randint(0, n, (8, 3), generator=g, device=dev)
    x = torch.view_as_complex(torch.rand((N, 2), generator=g, device=dev,
                                         dtype=torch.float64) * 2 - 1).reshape(n, n, n)
    idx = torch.arange(n, device=dev, dtype=torch.float64)
    for k in waves.tolist():
        ph = (k[0] * idx[:, None, None] + k[1] * idx[None, :, None] + k[2] * idx[None, None, :])
        x += torch.exp(1j * (2 * math.pi / n) * ph)
        del ph
    L = A.load()
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    shape = (C.c_int * 3)(n, n, n)
    flag = {"estimate": A.PFX_PLAN_ESTIMATE, "measure": A.PFX_PLAN_MEASURE,
            "patient": A.PFX_PLAN_PATIENT}[args.plan]
    out, y64 = {}, None
    for prec in ("f64", "f32"):
        code = A.PFX_F64 if prec == "f64" else A.PFX_F32
        buf = x.clone() if prec == "f64" else x.to(torch.complex64)
        nd = C.c_void_p()
        A.check(L.pfx_ndplan_create(C.byref(nd), code, A.PFX_C2C, 3, shape), "pfx_ndplan_create")
        if flag != A.PFX_PLAN_ESTIMATE:
            scratch = torch.empty_like(buf)
            A.check(L.pfx_ndplan_tune(nd.value, flag, scratch.data_ptr(), sp), "pfx_ndplan_tune")
            torch.cuda.synchronize()
            del scratch
        work = buf.clone()
        for _ in range(args.warmup):
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_FORWARD, work.data_ptr(), 1.0, sp), "fwd")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_FORWARD, work.data_ptr(), 1.0, sp), "fwd")
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        work.copy_(buf)
        A.check(L.pfx_ndplan_execute(nd.value, A.PFX_FORWARD, work.data_ptr(), 1.0, sp), "fwd")
        torch.cuda.synchronize()
        r = {"ms_per_forward": round(ms, 4),
             "tbs_algorithmic": round(2 * 3 * N * buf.element_size() / (ms * 1e-3) / 1e12, 3),
             "gflops": round(5.0 * N * math.log2(N) / (ms * 1e-3) / 1e9, 1)}
        if prec == "f64":
            y64 = work.clone()
            r["parseval_rel_err"] = float(abs(torch.linalg.vector_norm(work) ** 2 /
                                              (N * torch.linalg.vector_norm(buf) ** 2) - 1))
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_BACKWARD, work.data_ptr(), 1.0 / N, sp),
                    "bwd")
            torch.cuda.synchronize()
            r["roundtrip_rel_l2"] = float(torch.linalg.vector_norm(work - buf) /
                                          torch.linalg.vector_norm(buf))
        else:
            d = work.to(torch.complex128) - y64
            r["rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(y64))
            r["max_abs_rel_to_peak"] = float(d.abs().max() / y64.abs().max())
            del d
        log(f"[bench] 512^3 {prec}: {ms:.3f} ms")
        L.pfx_ndplan_destroy(nd.value)
        out[prec] = r
        del buf, work
    del x, y64
    torch.cuda.empty_cache()
    return {"workload": "3D C2C forward, in place, 512^3, U[-1,1) + 8 plane waves, one GPU",
            "results": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--precision", default="f32")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plan", default="measure", choices=["estimate", "measure", "patient"],
                    help="planning effort (FFTPlanFlag); measure/patient time the candidate "
                         "kernels on the data buffer before the input is generated")
    ap.add_argument("--workload", default="cube", choices=["cube", "pencil", "batch"],
                    help="cube: the 1024^3 C2C headline; pencil: one rank's share of a "
                         "2048^3 R2C on a 4x4 grid with on-device pack/unpack")
    ap.add_argument("--no-reference-driver", action="store_true",
                    help="batch workload: skip the one-ParaFaFT-object-per-shape baseline")
    ap.add_argument("--unfused", action="store_true",
                    help="pencil workload: separate pack / self copy / unpack launches")
    ap.add_argument("--no-graph", action="store_true",
                    help="cube workload: skip the HIP-graph replay (profiling runs keep only "
                         "the timed region's dispatches in the trace)")
    ap.add_argument("--no-split", action="store_true",
                    help="cube workload: skip the split-precision forms timed after plain FP32")
    ap.add_argument("--tune-cache-save", default=None,
                    help="write the tuning cache (every plan's chosen kernel) to this file at "
                         "the end, so a profiling run can replay the same kernels")
    ap.add_argument("--tune-cache-load", default=None,
                    help="load a tuning cache before planning (Measure then takes the cached "
                         "kernels without timing)")
    ap.add_argument("--no-extra", action="store_true",
                    help="cube workload: skip the 512^3, pencil-share and uneven-batch cases "
                         "reported beside the headline")
    ap.add_argument("--schedule", default="tuned", choices=["tuned", "chunked", "split"],
                    help="cube workload: the N-D schedule; tuned = the planner's choice, "
                         "chunked = no split axis 0, split = the fastest split axis 0")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of 4 GiB (the library's 2D-copy "
                         "kernel) five times, to scale rocprof FETCH/WRITE_SIZE counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {args.gpus}); the "
              "multi-GPU exchange is not part of it.", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    import torch
    import parafaft_amd as P
    from parafaft_amd import _capi as A

    L = A.load()
    if args.tune_cache_load:
        A.check(L.pfx_tune_cache_load(args.tune_cache_load.encode()), "pfx_tune_cache_load")
    if args.calibrate_copy:
        nbytes = 4 << 30
        a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(5):
            A.check(L.pfx_memcpy2d_async(b.data_ptr(), nbytes, a.data_ptr(), nbytes, nbytes, 1,
                                         torch.cuda.current_stream().cuda_stream), "copy")
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": nbytes, "launches": 5}), flush=True)
        return

    if args.workload in ("pencil", "batch"):
        r = (bench_pencil if args.workload == "pencil" else bench_batch)(args, torch, A)
        if args.tune_cache_save:
            A.check(L.pfx_tune_cache_save(args.tune_cache_save.encode()), "pfx_tune_cache_save")
        print(json.dumps(r), flush=True)
        return

    n = args.n
    D = 3
    N = n ** D
    prec = P.backend.precision_code(args.precision)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    esz_c = 16 if prec == A.PFX_F64 else 8
    log(f"[bench] allocating {N * esz_c / 2**30:.1f} GiB "
        f"complex<{'double' if prec == A.PFX_F64 else 'float'}> {n}^3")
    g = torch.Generator(device=dev)
    g.manual_seed(20241203)
    rdt = torch.float64 if prec == A.PFX_F64 else torch.float32
    x = (torch.rand((N, 2), device=dev, generator=g, dtype=rdt) * 2 - 1)
    x = torch.view_as_complex(x).reshape(n, n, n).contiguous()
    esz = x.element_size()

    nd = C.c_void_p()
    shape = (C.c_int * 3)(n, n, n)
    A.check(L.pfx_ndplan_create(C.byref(nd), prec, A.PFX_C2C, 3, shape), "pfx_ndplan_create")
    flag = {"estimate": A.PFX_PLAN_ESTIMATE, "measure": A.PFX_PLAN_MEASURE,
            "patient": A.PFX_PLAN_PATIENT}[args.plan]
    if args.schedule != "tuned":
        os.environ["PFX_ND_SCHEDULE"] = args.schedule
    if flag != A.PFX_PLAN_ESTIMATE:
        log(f"[bench] planning ({args.plan}) on the data buffer")
        A.check(L.pfx_ndplan_tune(nd.value, flag, x.data_ptr(), sp), "pfx_ndplan_tune")
        torch.cuda.synchronize()
        g.manual_seed(20241203)   # planning destroyed the buffer: regenerate the input
        xr = torch.view_as_real(x.view(-1))
        xr.copy_(torch.rand(xr.shape, device=dev, generator=g, dtype=rdt) * 2 - 1)
        torch.cuda.synchronize()
    if args.schedule != "tuned":
        os.environ.pop("PFX_ND_SCHEDULE", None)  # the headline plan only
    # A split axis 0 (where the planner finds it faster, e.g. 2048^3, or
    # forced by --schedule split) runs as whole forwards; its q-point pass is
    # timed alone for the roofline.  Otherwise the timed steps run the chunked
    # (or stage-by-stage) schedule part by part from its stage plans.
    split_c = C.c_int()
    A.check(L.pfx_ndplan_split(nd.value, C.byref(split_c)), "pfx_ndplan_split")
    split_g = split_c.value
    split_h = None
    if split_g:
        h = C.c_void_p()
        A.check(L.pfx_ndplan_split_stage(nd.value, C.byref(h)), "pfx_ndplan_split_stage")
        split_h = h.value
        log(f"[bench] split axis 0: G = {split_g}")
    stages = []
    for s in range(D):
        h = C.c_void_p()
        A.check(L.pfx_ndplan_stage(nd, s, C.byref(h)), "pfx_ndplan_stage")
        stages.append(h.value)
    def describe(h):
        buf = C.create_string_buffer(256)
        pb, th = C.c_int(), C.c_int()
        L.pfx_plan_describe(h, buf, 256, C.byref(pb), C.byref(th))
        return f"{buf.value.decode()}/tile{pb.value}/thr{th.value}"
    names = [describe(h) for h in stages]
    # chunked schedule: axes 2 and 1 run on `chunk` axis-0 planes at a time
    chunk_c = C.c_int()
    A.check(L.pfx_ndplan_chunk(nd.value, C.byref(chunk_c)), "pfx_ndplan_chunk")
    chunk = chunk_c.value
    cstages = []
    for s in range(D - 1 if chunk else 0):
        h = C.c_void_p()
        A.check(L.pfx_ndplan_chunk_stage(nd, s, C.byref(h)), "pfx_ndplan_chunk_stage")
        cstages.append(h.value)
    chunk_bytes = chunk * n * n * esz
    log(f"[bench] schedule: " + (f"axes 2+1 in chunks of {chunk} planes, then axis 0"
                                 if chunk else "stage by stage")
        + (f", axis 0 split ({split_g}-point DFTs fused into the row pass)" if split_g else ""))

    # ---- accuracy (outside the timed region): FP32 vs FP64 on the same input
    def accuracy(ndh):
        log("[bench] accuracy check vs FP64 path")
        x64 = x.to(torch.complex128)
        nd64 = C.c_void_p()
        A.check(L.pfx_ndplan_create(C.byref(nd64), A.PFX_F64, A.PFX_C2C, 3, shape), "nd64")
        y = x.clone()
        A.check(L.pfx_ndplan_execute(ndh, A.PFX_FORWARD, y.data_ptr(), 1.0, sp), "fwd32")
        A.check(L.pfx_ndplan_execute(nd64.value, A.PFX_FORWARD, x64.data_ptr(), 1.0, sp), "fwd64")
        torch.cuda.synchronize()
        diff = (y.to(torch.complex128) - x64)
        r = {"rel_l2_vs_fp64": float(torch.linalg.vector_norm(diff) /
                                     torch.linalg.vector_norm(x64)),
             "max_abs_vs_fp64": float(diff.abs().max())}
        r["max_abs_rel_to_peak"] = r["max_abs_vs_fp64"] / float(x64.abs().max())
        del diff, y, x64
        L.pfx_ndplan_destroy(nd64.value)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        return r

    def regenerate():
        g.manual_seed(20241203)
        xr = torch.view_as_real(x.view(-1))
        xr.copy_(torch.rand(xr.shape, device=dev, generator=g, dtype=rdt) * 2 - 1)
        torch.cuda.synchronize()

    acc = {}
    if not args.no_check and prec != A.PFX_F64:
        acc = accuracy(nd.value)

    # timed parts: the D stages, or (chunked) axes 2+1 together and axis 0
    parts = [[h] for h in stages] if not chunk else [None, [stages[D - 1]]]
    if split_g:
        parts = ["nd"]

    def step(events=None, st=None):
        st = st or sp
        base = x.data_ptr()
        for s, hs in enumerate(parts):
            if events is not None:
                events[s][0].record(stream)
            if hs == "nd":
                A.check(L.pfx_ndplan_execute(nd.value, A.PFX_FORWARD, base, 1.0, st), "fwd")
            elif hs is None:
                for c in range(n // chunk):
                    for h in cstages:
                        rc = L.pfx_execute(h, A.PFX_FORWARD, base + c * chunk_bytes, st)
                        if rc:
                            A.check(rc, "pfx_execute")
            else:
                for h in hs:
                    rc = L.pfx_execute(h, A.PFX_FORWARD, base, st)
                    if rc:
                        A.check(rc, "pfx_execute")
            if events is not None:
                events[s][1].record(stream)

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: exactly K forward transforms, synchronised on both sides
    NP = len(parts)
    ev = [[[torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)]
           for _ in range(NP)] for _ in range(args.steps)]
    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(ev[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eager_ms = (t1 - t0) * 1e3 / args.steps
    part_ms = [sum(ev[k][s][0].elapsed_time(ev[k][s][1]) for k in range(args.steps)) / args.steps
               for s in range(NP)]
    # per-step event time, first part's start to last part's end, as a median
    step_ev = sorted(ev[k][0][0].elapsed_time(ev[k][NP - 1][1]) for k in range(args.steps))
    median_ms = step_ev[len(step_ev) // 2] if len(step_ev) % 2 else \
        0.5 * (step_ev[len(step_ev) // 2 - 1] + step_ev[len(step_ev) // 2])
    log(f"[bench] eager ms/step {eager_ms:.3f} (median of events {median_ms:.3f}), "
        f"part ms {['%.3f' % t for t in part_ms]}")

    # The headline's timed region: the same forward captured once as a HIP
    # graph (one submission instead of one launch per chunk and stage) and
    # replayed exactly K times, synchronised on both sides, after one untimed
    # replay.  Without a graph the eager timing above is the value.
    graph_ms = None
    try:
        if args.no_graph:
            raise RuntimeError("skipped (--no-graph)")
        gstream = torch.cuda.Stream()      # the legacy null stream cannot be captured
        gsp = gstream.cuda_stream
        gh = C.c_void_p()
        torch.cuda.synchronize()
        A.check(L.pfx_graph_begin_capture(gsp), "pfx_graph_begin_capture")
        try:
            step(None, gsp)
        finally:
            A.check(L.pfx_graph_end_capture(gsp, C.byref(gh)), "pfx_graph_end_capture")
        A.check(L.pfx_graph_launch(gh.value, gsp), "pfx_graph_launch")
        torch.cuda.synchronize()
        g0 = time.perf_counter()
        for _ in range(args.steps):
            A.check(L.pfx_graph_launch(gh.value, gsp), "pfx_graph_launch")
        torch.cuda.synchronize()
        graph_ms = (time.perf_counter() - g0) * 1e3 / args.steps
        L.pfx_graph_destroy(gh.value)
        log(f"[bench] graph replay ms/step {graph_ms:.3f}")
    except Exception as e:  # graphs are reported, not required
        log(f"[bench] graph replay unavailable: {e}")
    ms_per_step = graph_ms if graph_ms else eager_ms
    timed_region = ("K replays of the forward captured as one HIP graph" if graph_ms else
                    "K eager forwards (one launch per chunk and stage)")

    # forward then backward with the 1/N scale fused into the last pass (the
    # reference benchmark's timed pair, bench_c2c.cpp:61-76), K pairs
    fb_ms = None
    if not args.no_graph:
        inv = 1.0 / N
        for _ in range(args.warmup):
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_FORWARD, x.data_ptr(), 1.0, sp), "fwd")
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_BACKWARD, x.data_ptr(), inv, sp), "bwd")
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for _ in range(args.steps):
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_FORWARD, x.data_ptr(), 1.0, sp), "fwd")
            A.check(L.pfx_ndplan_execute(nd.value, A.PFX_BACKWARD, x.data_ptr(), inv, sp), "bwd")
        torch.cuda.synchronize()
        fb_ms = round((time.perf_counter() - f0) * 1e3 / args.steps, 4)
        log(f"[bench] forward+backward ms {fb_ms}")

    flops = 5.0 * N * math.log2(N)
    gflops = flops / (ms_per_step * 1e-3) / 1e9
    alg_bytes = 2.0 * D * N * esz
    tbs = alg_bytes / (ms_per_step * 1e-3) / 1e12
    # dominant kernel: the single launch with the most time (axis 0 when the
    # other axes run in chunks)
    dom_name = None
    if split_g:
        # the q-point pass alone, in place on the array (same kernel and
        # geometry as from the scratch), K launches between two events
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        A.check(L.pfx_execute(split_h, A.PFX_FORWARD, x.data_ptr(), sp), "q-point pass")
        e0.record(stream)
        for _ in range(args.steps):
            A.check(L.pfx_execute(split_h, A.PFX_FORWARD, x.data_ptr(), sp), "q-point pass")
        e1.record(stream)
        torch.cuda.synchronize()
        dom, dom_ms = D - 1, e0.elapsed_time(e1) / args.steps
        dom_name = describe(split_h)
        part_ms = [part_ms[0], dom_ms]
    elif chunk:
        dom, dom_ms = D - 1, part_ms[1]
    else:
        dom = max(range(D), key=lambda s: part_ms[s])
        dom_ms = part_ms[dom]
    per_launch = 2.0 * N * esz
    achieved = per_launch / (dom_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_1024cube_f32.json")
    # the committed counter profile covers the 1024^3 FP32 configuration only
    if os.path.exists(tpath) and n == 1024 and prec == A.PFX_F32:
        try:
            with open(tpath) as f:
                t = json.load(f)
            st = t.get("stages", {}).get(str(dom), {})
            # counters of exactly the kernel this run times; otherwise none
            if st.get("plan_kernel") == (dom_name or names[dom]):
                traffic = st.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle sample)")
        try:
            cpu = cpu_baseline(n)
        except Exception as e:  # the oracle is optional on a box without gcc
            cpu = {"value": None, "unit": "GFLOP/s", "cores": 0, "kind": "port",
                   "sample": f"oracle unavailable: {e}"}

    out = {
        "metric": METRIC,
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {A.PFX_F32: "fp32", A.PFX_F64: "fp64"}.get(prec, args.precision),
        "data": "synthetic: iid Uniform[-1,1) re/im (seeded), resident in HBM",
        "config": {"workload": f"3D C2C forward, in place, {n}^3 "
                               f"complex<{'double' if prec == A.PFX_F64 else 'float'}>, one GPU "
                               "(one-rank fast path: "
                               + (f"axes 2 and 1 in chunks of {chunk} planes through the "
                                  "Infinity Cache, then axis 0)" if chunk and not split_g else
                                  f"axis 0 split as {split_g} x {n // max(split_g, 1)}: "
                                  f"{split_g}-point DFTs fused into the row pass of each chunk "
                                  "of planes, middle axis on the chunk, then the "
                                  f"{n // max(split_g, 1)}-point pass)" if split_g else
                                  "one launch per axis)"),
                   "grid": [n, n, n], "precision": args.precision, "planning": args.plan,
                   "chunk_planes": chunk,
                   "schedule": args.schedule,
                   "split_axis0": split_g,
                   "stage_kernels": names,
                   "chunk_kernels": [describe(h) for h in cstages]},
        "hbm_tbs_algorithmic": round(tbs, 3),
        "timed_region": timed_region,
        "eager_ms_per_step": round(eager_ms, 4),
        "event_median_ms_per_step": round(median_ms, 4),
        "graph_replay_ms_per_step": round(graph_ms, 4) if graph_ms else None,
        "forward_backward_ms": fb_ms,
        "part_ms": ({"forward (split axis 0)": round(part_ms[0], 4),
                     "q-point pass (alone)": round(part_ms[1], 4)} if split_g else
                    {"axes 2+1 (chunked)": round(part_ms[0], 4), "axis 0": round(part_ms[1], 4)}
                    if chunk else {f"axis {D - 1 - s}": round(t, 4) for s, t in enumerate(part_ms)}),
        "accuracy": acc,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel": dom_name or names[dom],
                     "algorithmic_bytes_per_launch": per_launch},
        "cpu_baseline": cpu,
    }
    L.pfx_ndplan_destroy(nd.value)

    def time_parts(shs, scs, pchunk):
        """Event-timed parts of an N-D forward given its stage handles (and
        chunk-stage handles): axes 2+1 in chunks and axis 0, or each axis;
        the mean over K steps."""
        if pchunk:
            pparts = [("axes 2+1 (chunked)", None), ("axis 0", [shs[D - 1]])]
        else:
            pparts = [(f"axis {D - 1 - st}", [h]) for st, h in enumerate(shs)]
        cb = pchunk * n * n * esz
        acc_ms = {lbl: 0.0 for lbl, _ in pparts}
        for k in range(args.steps + 1):
            evs = []
            for lbl, hs in pparts:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                if hs is None:
                    for c in range(n // pchunk):
                        for h in scs:
                            A.check(L.pfx_execute(h, A.PFX_FORWARD, x.data_ptr() + c * cb, sp),
                                    "pfx_execute")
                else:
                    for h in hs:
                        A.check(L.pfx_execute(h, A.PFX_FORWARD, x.data_ptr(), sp), "pfx_execute")
                e1.record(stream)
                evs.append((lbl, e0, e1))
            torch.cuda.synchronize()
            if k:  # the first pass is a warm-up
                for lbl, e0, e1 in evs:
                    acc_ms[lbl] += e0.elapsed_time(e1) / args.steps
        return {lbl: round(v, 4) for lbl, v in acc_ms.items()}

    # ---- split-precision MFMA forms on the same grid and input (the metric's
    # "split vs plain"): planned the same way, checked against the same FP64
    # path, timed over the same K forwards
    if prec == A.PFX_F32 and not args.no_split and not split_g:
        split = {}
        for mode in ("f32_split_f16", "f32_split_bf16"):
            nds = C.c_void_p()
            A.check(L.pfx_ndplan_create(C.byref(nds), P.backend.precision_code(mode), A.PFX_C2C,
                                        3, shape), "pfx_ndplan_create")
            if flag != A.PFX_PLAN_ESTIMATE:
                log(f"[bench] planning {mode} ({args.plan})")
                A.check(L.pfx_ndplan_tune(nds.value, flag, x.data_ptr(), sp), "pfx_ndplan_tune")
                torch.cuda.synchronize()
                regenerate()
            sacc = accuracy(nds.value) if not args.no_check else {}
            for _ in range(args.warmup):
                A.check(L.pfx_ndplan_execute(nds.value, A.PFX_FORWARD, x.data_ptr(), 1.0, sp), "fwd")
            torch.cuda.synchronize()
            s0 = time.perf_counter()
            for _ in range(args.steps):
                A.check(L.pfx_ndplan_execute(nds.value, A.PFX_FORWARD, x.data_ptr(), 1.0, sp), "fwd")
            torch.cuda.synchronize()
            sms = (time.perf_counter() - s0) * 1e3 / args.steps
            schunk = C.c_int()
            A.check(L.pfx_ndplan_chunk(nds.value, C.byref(schunk)), "pfx_ndplan_chunk")
            kern, shs = [], []
            for st in range(D):
                h = C.c_void_p()
                A.check(L.pfx_ndplan_stage(nds, st, C.byref(h)), "pfx_ndplan_stage")
                kern.append(describe(h.value))
                shs.append(h.value)
            scs = []
            for st in range(D - 1 if schunk.value else 0):
                h = C.c_void_p()
                A.check(L.pfx_ndplan_chunk_stage(nds, st, C.byref(h)), "pfx_ndplan_chunk_stage")
                scs.append(h.value)
            sparts = time_parts(shs, scs, schunk.value)
            split[mode] = {"ms_per_step": round(sms, 4),
                           "gflops": round(flops / (sms * 1e-3) / 1e9, 2),
                           "speed_vs_plain": round(eager_ms / sms, 4),
                           "accuracy": sacc, "chunk_planes": schunk.value,
                           "stage_kernels": kern,
                           "part_ms": sparts}
            log(f"[bench] {mode}: {sms:.3f} ms/step ({eager_ms / sms:.3f}x plain, both eager)")
            L.pfx_ndplan_destroy(nds.value)
        # What caps the split forms, from the parts measured above: both
        # forms move the same bytes.  FP16's column passes run radix-32
        # sub-DFTs on the 32x32 matrix cores over 32-column tiles (256-byte
        # runs, as plain); BF16's three-term operands do not fit those
        # registers, so its column kernels keep 16 columns (128-byte runs).
        # The split/scale work around the matrix-core products adds VALU
        # and matrix-core time to every pass (most for BF16's six products).
        plain_parts = out["part_ms"]
        split["ceiling"] = {
            "plain_part_ms": plain_parts,
            "split_minus_plain_ms": {m: {k: round(v["part_ms"].get(k, 0.0) - t, 4)
                                         for k, t in plain_parts.items()}
                                     for m, v in split.items()},
            "note": ("same bytes, both HBM-bound: FP16's column passes run radix-32 "
                     "matrix-core sub-DFTs over 32-column (256-byte-run) tiles like plain, "
                     "BF16's keep 16-column (128-byte) tiles; every pass adds the split, "
                     "scale and matrix-core work around the products; plain FP32 sits at "
                     "the measured HBM touch ceiling, so no arithmetic form can pass it")}
        out["split_vs_plain"] = split
    # the other measured cases on the same run (the driver records only this
    # line): 512^3 FP64/FP32, one rank's R2C pencil share of 2048^3 on 4x4,
    # and the 512 uneven transforms
    if n == 1024 and prec == A.PFX_F32 and not args.no_extra:
        import copy
        del x
        torch.cuda.empty_cache()
        extra = {}
        xargs = copy.copy(args)
        xargs.no_reference_driver = True   # its 512 graphs alone take longer than the rest
        for label, fn in (("512cube", bench_512), ("pencil_share", bench_pencil),
                          ("uneven_batch", bench_batch)):
            t0 = time.perf_counter()
            try:
                extra[label] = fn(xargs, torch, A)
            except Exception as e:  # reported, the headline stands
                extra[label] = {"error": f"{type(e).__name__}: {e}"}
            log(f"[bench] {label}: {time.perf_counter() - t0:.1f} s")
            torch.cuda.empty_cache()
        out["configs"] = extra
    # the metric's accuracy bound (rel-L2 <= 2e-6 against FP64) for every
    # FP32 form measured; a run that breaks it exits non-zero after printing
    ok = None
    if acc:
        forms = [acc] + [v["accuracy"] for v in out.get("split_vs_plain", {}).values()
                         if v.get("accuracy")]
        ok = all(f["rel_l2_vs_fp64"] <= ACC_BOUND for f in forms)
    out["accuracy_ok"] = ok
    if args.tune_cache_save:
        A.check(L.pfx_tune_cache_save(args.tune_cache_save.encode()), "pfx_tune_cache_save")
    print(json.dumps(out), flush=True)
    if ok is False:
        log(f"[bench] accuracy bound rel-L2 <= {ACC_BOUND} vs FP64 exceeded")
        sys.exit(3)


if __name__ == "__main__":
    main()
