"""One rank's pencil share of ParaFaFT_R2C<D>::forward_in_place on one GPU.

Restates parafaft_r2c.hpp:360-430 for the rank at `coords` of a (D-1)-dim
process grid `dims`, with the packed exchange of exchange_packed
(parafaft_common.hpp:564-669) kept on the device:

  stage 0   R2C along the last axis of the padded real share
  for t in 0..D-2 (grid dimension D-2-t, nparts = dims[D-2-t]):
     pack src -> send segments            (pfx_pack, bit-exact wire order)
     self block: send segment -> recv segment (the self-copy, :639-646)
     peer blocks: recv segments already resident on the device
     unpack recv segments -> dst          (pfx_unpack)
     C2C along axis D-2-t on dst
     swap(src, dst)

Every transition has nparts send and nparts receive segments, all
device-resident: for a 3D share on a 4 x 4 grid that is 16 peer buffers.
The peers' incoming blocks are whatever the caller places in the receive
segments (`recv_segment`); nothing crosses to another GPU.

Fused form (`fused=True`, the default where it applies): the pack and the
unpack move into the transforms (pfx_execute_segmented).  Stage t's last
pass stores its output straight into transition t's segments, the rank's
own block straight into its receive segment (so no self copy), and stage
t + 1's first pass loads from transition t's receive segments.  The
segments hold bit-identical data to the unfused path; only the HBM round
trips of pack, self copy and unpack are gone.
"""
from __future__ import annotations

import ctypes as C
import math

from . import _capi as A
from .backend import DeviceBuffer, Plan, precision_code, ptr
from .geometry import exchange_geometry, r2c_stage_shapes


def _fusable(plan):
    name = plan.describe()[0]
    return "direct_dft" not in name and "split" not in name


class PencilShareR2C:
    def __init__(self, global_real_shape, dims, coords, precision="f32", fast=True, stream=None,
                 fused=None):
        self.D = D = len(global_real_shape)
        self.global_real = [int(v) for v in global_real_shape]
        self.dims = list(dims)
        self.coords = list(coords)
        self.precision = precision_code(precision)
        self.esz = 16 if self.precision == A.PFX_F64 else 8
        self.fast = fast
        self.stream = ptr(stream)
        ins, outs, self.real_start, self.complex_start = r2c_stage_shapes(
            self.global_real, self.dims, self.coords)
        self.stage_shapes, self.stage_output_shapes = ins, outs
        self.gc = list(self.global_real)
        self.gc[-1] = self.global_real[-1] // 2 + 1
        self.max_complex = max(math.prod(s) for s in outs)
        # transitions
        self.nparts, self.me, self.geoms = [], [], []
        for t in range(D - 1):
            gd = D - 2 - t
            self.nparts.append(self.dims[gd])
            self.me.append(self.coords[gd])
            self.geoms.append(exchange_geometry(self.dims[gd], outs[t], D - 1 - t, ins[t + 1],
                                                D - 2 - t))
        self.send = [DeviceBuffer(max(1, sum(g.send_counts)) * self.esz) for g in self.geoms]
        self.recv = [DeviceBuffer(max(1, sum(g.recv_counts)) * self.esz) for g in self.geoms]
        # plans: stage 0 R2C, stages 1.. C2C (create_backend_plans, r2c:1158-1206)
        n = self.global_real[-1]
        batch0 = math.prod(ins[0][:D - 1])
        self.r2c = Plan.r2c(self.precision, n, batch0, 1, 2 * self.gc[-1])
        self.plans = [None]
        self.leading = [0]
        self.slice = [0]
        for stage in range(1, D):
            axis = D - 1 - stage
            length = self.gc[axis]
            shp = ins[stage]
            # (strides only matter when there are transforms; an empty share,
            # where decompose gives this rank no elements, has none)
            if axis == 0:
                batch = math.prod(shp[1:])
                self.plans.append(Plan.c2c(self.precision, length, batch, max(1, batch), 1))
                self.leading.append(1)
                self.slice.append(0)
            else:
                trailing = math.prod(shp[axis + 1:])
                leading = math.prod(shp[:axis])
                if fast:
                    self.plans.append(Plan.c2c(self.precision, length, trailing, max(1, trailing),
                                               1, leading, length * trailing))
                    self.leading.append(1)
                else:
                    self.plans.append(Plan.c2c(self.precision, length, trailing, max(1, trailing),
                                               1))
                    self.leading.append(leading)
                self.slice.append(length * trailing * self.esz)
        # the fused form needs one launch per stage (fast plans), tiled
        # plain-precision kernels, and segment maps the kernels take (at most
        # 16 peers per grid dimension)
        tiled = all(_fusable(pl) for pl in [self.r2c] + self.plans[1:])
        can_fuse = fast and tiled and all(1 <= n <= 16 for n in self.nparts)
        self.fused = can_fuse if fused is None else bool(fused)
        if self.fused and not can_fuse:
            raise ValueError("the fused exchange needs fast=True, tiled FP32/FP64 kernels and "
                             "at most 16 parts per grid dimension")
        self.scratch = None if self.fused else DeviceBuffer(self.max_complex * self.esz)

    # sizes ---------------------------------------------------------------
    def get_required_output_size(self):
        """Scalars the padded buffer must hold (parafaft_r2c.hpp:573-596)."""
        stage0 = math.prod(self.stage_shapes[0][:self.D - 1]) * 2 * self.gc[-1]
        return max(stage0, 2 * self.max_complex)

    def recv_segment(self, t: int, p: int):
        """(device address, element count) of the block peer p sends in
        transition t."""
        g = self.geoms[t]
        return self.recv[t].ptr + g.recv_displs[p] * self.esz, g.recv_counts[p]

    def send_segment(self, t: int, p: int):
        g = self.geoms[t]
        return self.send[t].ptr + g.send_displs[p] * self.esz, g.send_counts[p]

    # stages --------------------------------------------------------------
    def _arr(self, ctype, vals):
        return (ctype * len(vals))(*vals)

    def pack(self, t, src):
        g = self.geoms[t]
        A.check(A.load().pfx_pack(self.send[t].ptr, ptr(src), len(g.src_n),
                                  self._arr(C.c_int, g.src_n), self._arr(C.c_int, g.src_s),
                                  self._arr(C.c_uint64, g.send_displs), g.src_leading,
                                  g.src_trailing, g.src_axis_extent, self.esz, self.stream),
                "pfx_pack")

    def self_copy(self, t):
        g = self.geoms[t]
        me = self.me[t]
        A.check(A.load().pfx_memcpy_async(self.recv[t].ptr + g.recv_displs[me] * self.esz,
                                          self.send[t].ptr + g.send_displs[me] * self.esz,
                                          g.send_counts[me] * self.esz, self.stream),
                "pfx_memcpy_async")

    def unpack(self, t, dst):
        g = self.geoms[t]
        A.check(A.load().pfx_unpack(ptr(dst), self.recv[t].ptr, len(g.dst_n),
                                    self._arr(C.c_int, g.dst_n), self._arr(C.c_int, g.dst_s),
                                    self._arr(C.c_uint64, g.recv_displs), g.dst_leading,
                                    g.dst_trailing, g.dst_axis_extent, self.esz, self.stream),
                "pfx_unpack")

    def tune(self, padded, reps: int = 2):
        """Measure planning for this share's schedule (the FFTW_MEASURE analogue
        for the whole forward, fused or stepwise): for each stage plan in turn,
        every candidate kernel is timed inside the full forward on `padded`
        (contents destroyed) and the fastest is kept.  Kernels that read or
        write segments carry extra addressing registers, so the best choice
        for a fused stage can differ from the best for the plain stage."""
        import time
        L = A.load()

        def timed():
            self.forward_in_place(padded)
            A.check(L.pfx_stream_synchronize(self.stream), "sync")
            t0 = time.perf_counter()
            for _ in range(reps):
                self.forward_in_place(padded)
            A.check(L.pfx_stream_synchronize(self.stream), "sync")
            return time.perf_counter() - t0

        chosen = []
        for plan in [self.r2c] + self.plans[1:]:
            best_i, best_t = None, None
            for i in range(plan.num_candidates()):
                try:
                    plan.use_candidate(i)
                    t = timed()
                except A.PfxError:
                    continue
                if best_t is None or t < best_t:
                    best_i, best_t = i, t
            if best_i is not None:
                plan.use_candidate(best_i)
            chosen.append(plan.describe()[0])
        return chosen

    def segmap(self, t, side):
        """pfx_segmap of transition t: side "out" is where stage t stores
        (send segments, the rank's own block in its receive segment), side
        "in" where stage t + 1 loads (receive segments)."""
        g, me = self.geoms[t], self.me[t]
        k = self.nparts[t]
        if side == "out":
            segs = [self.send[t].ptr + g.send_displs[p] * self.esz for p in range(k)]
            segs[me] = self.recv[t].ptr + g.recv_displs[me] * self.esz
            start, count, trailing = g.src_s, g.src_n, g.src_trailing
        else:
            segs = [self.recv[t].ptr + g.recv_displs[p] * self.esz for p in range(k)]
            start, count, trailing = g.dst_s, g.dst_n, g.dst_trailing
        arrays = (self._arr(C.c_void_p, segs), self._arr(C.c_int, start),
                  self._arr(C.c_int, count))
        m = A.SegMap(k, arrays[0], arrays[1], arrays[2], trailing)
        m._keep = arrays
        return m

    def _execute_segmented(self, plan, data, min_, mout):
        A.check(A.load().pfx_execute_segmented(
            plan.handle, A.PFX_FORWARD, data, C.byref(min_) if min_ is not None else None,
            C.byref(mout) if mout is not None else None, 1.0, self.stream),
            "pfx_execute_segmented")

    def forward_fused(self, padded, events=None):
        """Stage 0 R2C storing into transition 0's segments, then per stage a
        C2C loading from the previous transition's receive segments and
        storing into the next transition's segments (the last stage stores
        into `padded`)."""
        mark = events or (lambda label: None)
        mark("r2c")
        self._execute_segmented(self.r2c, ptr(padded), None, self.segmap(0, "out"))
        for t in range(self.D - 1):
            stage = t + 1
            mark(f"fft{stage}")
            mout = self.segmap(stage, "out") if stage < self.D - 1 else None
            self._execute_segmented(self.plans[stage], ptr(padded), self.segmap(t, "in"), mout)
        mark("end")

    def fft_stage(self, stage, buf):
        base = ptr(buf)
        for i in range(self.leading[stage]):
            self.plans[stage].execute(A.PFX_FORWARD, base + i * self.slice[stage], self.stream)

    def forward_in_place(self, padded, events=None):
        """Stage 0 R2C, then per transition pack / self copy / unpack / C2C
        (or the fused form, see forward_fused).  `events`, if given, is
        called with a label before each step (timing)."""
        if self.fused:
            return self.forward_fused(padded, events)
        mark = events or (lambda label: None)
        mark("r2c")
        self.r2c.execute(A.PFX_FORWARD, padded, self.stream)
        src, dst = ptr(padded), self.scratch.ptr
        if (self.D - 1) % 2 == 1:
            A.check(A.load().pfx_memcpy_async(self.scratch.ptr, ptr(padded),
                                              math.prod(self.stage_output_shapes[0]) * self.esz,
                                              self.stream), "pfx_memcpy_async")
            src, dst = self.scratch.ptr, ptr(padded)
        for t in range(self.D - 1):
            mark(f"pack{t}")
            self.pack(t, src)
            self.self_copy(t)
            mark(f"unpack{t}")
            self.unpack(t, dst)
            mark(f"fft{t + 1}")
            self.fft_stage(t + 1, dst)
            src, dst = dst, src
        mark("end")
