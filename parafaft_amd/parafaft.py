"""One-rank ParaFaFT drivers over MI355Backend.

``ParaFaFT_C2C`` / ``ParaFaFT_R2C`` restate, for a single rank, the stage loop
of the reference drivers so the backend is exercised exactly as the reference
drives it:

* plan geometry per stage -- ParaFaFT_C2C::create_backend_plans
  (parafaft_c2c.hpp:249-287), ParaFaFT_R2C::create_backend_plans
  (parafaft_r2c.hpp:1158-1206);
* execution per stage -- perform_fft_on_buffer (parafaft_c2c.hpp:866-897,
  parafaft_r2c.hpp:1219-1245), including the per-slice loop on middle axes;
* the ping-pong buffer parity and the one-rank exchange, which is an identity
  flat copy (exchange_local, parafaft_common.hpp:467-543);
* shape accessors (parafaft_c2c.hpp:348-426, parafaft_r2c.hpp:552-796).

``fast=True`` selects the one-GPU fast path instead: one launch per axis via a
pfx_ndplan, no exchange copies and no per-slice launches; the results are
identical because on one rank every exchange is the identity.
"""
from __future__ import annotations

import ctypes as C
import math

from . import _capi as A
from .backend import DeviceBuffer, FFTDirection, FFTPlanFlag, MI355Backend, precision_code, ptr


def _prod(v):
    return int(math.prod(v)) if v else 1


class _NdPlan:
    def __init__(self, precision, kind, shape):
        L = A.load()
        h = C.c_void_p()
        arr = (C.c_int * len(shape))(*shape)
        A.check(L.pfx_ndplan_create(C.byref(h), precision, kind, len(shape), arr),
                "pfx_ndplan_create")
        self.handle = h.value

    def execute(self, direction, data, scale, stream):
        A.check(A.load().pfx_ndplan_execute(self.handle, direction, ptr(data), scale, stream),
                "pfx_ndplan_execute")

    def execute_r2c_oop(self, direction, real, padded, scale, stream) -> bool:
        """Out-of-place R2C/C2R with the padding copy fused into the real-axis
        pass; False when the plan has no tiled kernel for it."""
        rc = A.load().pfx_ndplan_execute_r2c_oop(self.handle, direction, ptr(real), ptr(padded),
                                                  scale, stream)
        if rc == A.PFX_ERR_UNSUPPORTED:
            return False
        A.check(rc, "pfx_ndplan_execute_r2c_oop")
        return True

    def plan(self, flag: FFTPlanFlag, nbytes: int, stream):
        """Measure / Patient planning (kernels and chunked or stage-by-stage
        schedule) on a temporary device buffer of the array's size, freed
        again: the user's arrays are never touched."""
        if flag == FFTPlanFlag.Estimate:
            return
        tmp = DeviceBuffer(nbytes)
        L = A.load()
        rc = L.pfx_ndplan_tune(self.handle, flag.value, tmp.ptr, stream)
        L.pfx_stream_synchronize(stream)
        del tmp
        A.check(rc, "pfx_ndplan_tune")

    def __del__(self):
        try:
            if self.handle:
                A.load().pfx_ndplan_destroy(self.handle)
                self.handle = 0
        except Exception:
            pass


class ParaFaFT_C2C:
    """Single-rank ParaFaFT_C2C<D, MI355Backend> (parafaft_c2c.hpp:69-1108)."""

    def __init__(self, global_shape, precision="f64",
                 plan_flag: FFTPlanFlag = FFTPlanFlag.Estimate, fast: bool = False,
                 stream=None):
        self.D = D = len(global_shape)
        self.global_shape = [int(n) for n in global_shape]
        self.precision = precision_code(precision)
        self.esz = 16 if self.precision == A.PFX_F64 else 8
        self.fast = fast
        self.backend = MI355Backend(D, plan_flag, self.precision, stream)
        # one rank: every stage shape is the global shape (setup_stage_shapes)
        self.stage_shapes = [list(self.global_shape) for _ in range(D)]
        self.max_stage_size = _prod(self.global_shape)
        if fast:
            self._nd = _NdPlan(self.precision, A.PFX_C2C, self.global_shape)
            self._nd.plan(plan_flag, self.max_stage_size * self.esz, self.backend.stream)
            self.scratch = None
        else:
            self.scratch = DeviceBuffer(self.max_stage_size * self.esz)
            self._create_backend_plans()

    def _create_backend_plans(self):
        D = self.D
        for stage in range(D):
            axis = D - 1 - stage
            length = self.global_shape[axis]
            shp = self.stage_shapes[stage]
            if axis == D - 1:
                self.backend.create_stage_plan(stage, length, _prod(shp[:D - 1]), self.scratch,
                                               1, length)
            elif axis == 0:
                batch = _prod(shp[1:])
                self.backend.create_stage_plan(stage, length, batch, self.scratch, batch, 1)
            else:
                trailing = _prod(shp[axis + 1:])
                self.backend.create_stage_plan(stage, length, trailing, self.scratch, trailing, 1)

    # accessors (parafaft_c2c.hpp:348-426)
    def get_local_size(self):
        return _prod(self.stage_shapes[0])

    def get_required_output_size(self):
        return self.max_stage_size

    def get_local_shape(self):
        return list(self.stage_shapes[0])

    def get_global_start(self):
        return [0] * self.D

    def get_final_shape(self):
        return list(self.stage_shapes[self.D - 1])

    def get_final_start(self):
        return [0] * self.D

    def get_domain_decomposition(self):
        """parafaft_c2c.hpp:453-458 (one rank)."""
        return [1] * self.D

    def get_grid_ndims(self):
        """parafaft_c2c.hpp:463: the grid has D - 1 dimensions."""
        return self.D - 1

    def get_grid_coords(self):
        """parafaft_c2c.hpp:473-477 (one rank: the origin)."""
        return [0] * (self.D - 1)

    def get_buffer(self):
        """parafaft_c2c.hpp:439-445: no backend-owned buffer here."""
        return None

    def _perform_fft_on_buffer(self, stage, axis, direction, buf, scale=1.0):
        D = self.D
        if axis == D - 1 or axis == 0:
            self.backend.execute_stage(stage, direction, buf, scale)
        else:
            shp = self.stage_shapes[stage]
            leading = _prod(shp[:axis])
            trailing = _prod(shp[axis + 1:])
            slice_bytes = self.global_shape[axis] * trailing * self.esz
            base = ptr(buf)
            for i in range(leading):
                self.backend.execute_stage(stage, direction, base + i * slice_bytes, scale)

    def _run(self, data, forward: bool, scale: float, sync: bool = True):
        D = self.D
        if self.fast:
            self._nd.execute(A.PFX_FORWARD if forward else A.PFX_BACKWARD, data, scale,
                             self.backend.stream)
            if sync:
                self.backend.sync()
            return
        nbytes = self.max_stage_size * self.esz
        if (D - 1) % 2 == 0:
            src, dst = ptr(data), self.scratch.ptr
        else:
            self.backend.memcpy(self.scratch.ptr, ptr(data), nbytes)
            src, dst = self.scratch.ptr, ptr(data)
        stages = range(D) if forward else range(D - 1, -1, -1)
        direction = FFTDirection.Forward if forward else FFTDirection.Backward
        last = D - 1 if forward else 0
        for stage in stages:
            axis = D - 1 - stage
            self._perform_fft_on_buffer(stage, axis, direction, src,
                                        scale if stage == last else 1.0)
            if stage != last:
                # exchange_local on one rank: identity flat copy
                self.backend.memcpy(dst, src, nbytes)
                src, dst = dst, src
        if sync:
            self.backend.sync()

    def forward(self, data, sync: bool = True):
        """parafaft_c2c.hpp:539-602 (unnormalised).  ``sync=False`` leaves the
        closing stream synchronisation (c2c:601) to the caller, e.g. while the
        stream is being captured into a graph."""
        self._run(data, True, 1.0, sync)

    def backward(self, data, scale: float = 1.0, sync: bool = True):
        """parafaft_c2c.hpp:632-694; ``scale`` (default 1, the reference's
        unnormalised convention) is fused into the last pass when != 1."""
        self._run(data, False, scale, sync)


class ParaFaFT_R2C:
    """Single-rank ParaFaFT_R2C<D, MI355Backend> (parafaft_r2c.hpp:98-1308),
    in-place padded form."""

    def __init__(self, global_real_shape, precision="f64",
                 plan_flag: FFTPlanFlag = FFTPlanFlag.Estimate, fast: bool = False,
                 stream=None):
        self.D = D = len(global_real_shape)
        self.global_real_shape = [int(n) for n in global_real_shape]
        self.global_complex_shape = list(self.global_real_shape)
        self.global_complex_shape[-1] = self.global_real_shape[-1] // 2 + 1
        self.precision = precision_code(precision)
        self.esz = 16 if self.precision == A.PFX_F64 else 8
        self.fast = fast
        self.backend = MI355Backend(D, plan_flag, self.precision, stream)
        self.stage_shapes = [list(self.global_real_shape)] + \
            [list(self.global_complex_shape) for _ in range(D - 1)]
        self.stage_output_shapes = [list(self.global_complex_shape) for _ in range(D)]
        self.max_complex_size = _prod(self.global_complex_shape)
        if fast:
            self._nd = _NdPlan(self.precision, A.PFX_R2C, self.global_real_shape)
            self._nd.plan(plan_flag, self.get_required_output_size() * self.esz // 2,
                          self.backend.stream)
        else:
            self.scratch = DeviceBuffer(self.max_complex_size * self.esz)
            self._create_backend_plans()

    def _create_backend_plans(self):
        D = self.D
        batch0 = _prod(self.stage_shapes[0][:D - 1])
        n = self.global_real_shape[-1]
        pd = 2 * self.global_complex_shape[-1]
        self.backend.create_r2c_inplace_plan(n, batch0, self.scratch, 1, pd)
        self.backend.create_c2r_inplace_plan(n, batch0, self.scratch, 1, pd)
        for stage in range(1, D):
            axis = D - 1 - stage
            length = self.global_complex_shape[axis]
            shp = self.stage_shapes[stage]
            if axis == 0:
                batch = _prod(shp[1:])
                self.backend.create_stage_plan(stage, length, batch, self.scratch, batch, 1)
            else:
                trailing = _prod(shp[axis + 1:])
                self.backend.create_stage_plan(stage, length, trailing, self.scratch, trailing, 1)

    def get_required_output_size(self):
        """parafaft_r2c.hpp:573-596, in scalars."""
        stage0 = _prod(self.stage_shapes[0][:self.D - 1]) * 2 * self.global_complex_shape[-1]
        return max(stage0, 2 * self.max_complex_size)

    def get_local_real_shape(self):
        return list(self.stage_shapes[0])

    def get_local_complex_shape(self):
        return list(self.stage_shapes[self.D - 1])

    def get_local_complex_size(self):
        return _prod(self.stage_shapes[self.D - 1])

    def get_real_global_start(self):
        return [0] * self.D

    def get_complex_global_start(self):
        return [0] * self.D

    def get_domain_decomposition(self):
        """parafaft_r2c.hpp:712-717 (one rank)."""
        return [1] * self.D

    def get_grid_ndims(self):
        return self.D - 1

    def get_grid_coords(self):
        return [0] * (self.D - 1)

    def get_real_buffer(self):
        """parafaft_r2c.hpp:773-790: no backend-owned buffers here."""
        return None

    def get_buffer(self):
        return None

    def _perform_fft_on_buffer(self, stage, axis, direction, buf, scale=1.0):
        if axis == 0:
            self.backend.execute_stage(stage, direction, buf, scale)
        else:
            shp = self.stage_shapes[stage]
            leading = _prod(shp[:axis])
            trailing = _prod(shp[axis + 1:])
            slice_bytes = self.global_complex_shape[axis] * trailing * self.esz
            base = ptr(buf)
            for i in range(leading):
                self.backend.execute_stage(stage, direction, base + i * slice_bytes, scale)

    def forward_in_place(self, padded, sync: bool = True):
        """parafaft_r2c.hpp:360-430 (``sync`` as in ParaFaFT_C2C.forward)."""
        D = self.D
        if self.fast:
            self._nd.execute(A.PFX_FORWARD, padded, 1.0, self.backend.stream)
            if sync:
                self.backend.sync()
            return
        nbytes = self.max_complex_size * self.esz
        self.backend.execute_r2c_inplace(padded)
        if (D - 1) % 2 == 0:
            src, dst = ptr(padded), self.scratch.ptr
        else:
            self.backend.memcpy(self.scratch.ptr, ptr(padded), nbytes)
            src, dst = self.scratch.ptr, ptr(padded)
        for stage in range(1, D):
            axis = D - 1 - stage
            self.backend.memcpy(dst, src, nbytes)  # exchange_local: identity
            self._perform_fft_on_buffer(stage, axis, FFTDirection.Forward, dst)
            src, dst = dst, src
        if sync:
            self.backend.sync()

    def copy_real_to_padded(self, real_input, padded_output):
        """parafaft_r2c.hpp:804-823: contiguous real rows into padded rows of
        2(N/2+1) scalars, one 2D copy."""
        n = self.global_real_shape[-1]
        rsz = self.esz // 2
        rows = _prod(self.stage_shapes[0][:self.D - 1])
        self.backend.memcpy2d(padded_output, 2 * self.global_complex_shape[-1] * rsz,
                              real_input, n * rsz, n * rsz, rows)

    def copy_padded_to_real(self, padded_input, real_output):
        """parafaft_r2c.hpp:825-847: padded rows back to contiguous real rows."""
        n = self.global_real_shape[-1]
        rsz = self.esz // 2
        rows = _prod(self.stage_shapes[0][:self.D - 1])
        self.backend.memcpy2d(real_output, n * rsz, padded_input,
                              2 * self.global_complex_shape[-1] * rsz, n * rsz, rows)

    def forward(self, real_input, complex_output, sync: bool = True):
        """parafaft_r2c.hpp:333-339: real input (unpadded, untouched) to the
        spectrum in `complex_output` (get_required_output_size() scalars, also
        the working space).  Fast mode fuses the padding copy into the R2C
        pass."""
        if self.fast and self._nd.execute_r2c_oop(A.PFX_FORWARD, real_input, complex_output,
                                                  1.0, self.backend.stream):
            if sync:
                self.backend.sync()
            return
        self.copy_real_to_padded(real_input, complex_output)
        self.forward_in_place(complex_output, sync)

    def backward(self, complex_input, real_output, scale: float = 1.0):
        """parafaft_r2c.hpp:446-453: the spectrum in `complex_input` (used as
        working space, destroyed) to the real output.  Fast mode fuses the
        padding copy into the C2R pass."""
        if self.fast and self._nd.execute_r2c_oop(A.PFX_BACKWARD, real_output, complex_input,
                                                  scale, self.backend.stream):
            self.backend.sync()
            return
        self.backward_in_place(complex_input, scale)
        self.copy_padded_to_real(complex_input, real_output)
        self.backend.sync()

    def backward_in_place(self, padded, scale: float = 1.0):
        """parafaft_r2c.hpp:473-542 (unnormalised unless scale != 1)."""
        D = self.D
        if self.fast:
            self._nd.execute(A.PFX_BACKWARD, padded, scale, self.backend.stream)
            self.backend.sync()
            return
        nbytes = self.max_complex_size * self.esz
        if (D - 1) % 2 == 0:
            src, dst = ptr(padded), self.scratch.ptr
        else:
            self.backend.memcpy(self.scratch.ptr, ptr(padded), nbytes)
            src, dst = self.scratch.ptr, ptr(padded)
        for stage in range(D - 1, 0, -1):
            axis = D - 1 - stage
            self._perform_fft_on_buffer(stage, axis, FFTDirection.Backward, src)
            self.backend.memcpy(dst, src, nbytes)
            src, dst = dst, src
        self.backend.execute_c2r_inplace(padded, scale)
        self.backend.sync()
