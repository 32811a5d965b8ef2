"""Host-side pencil geometry: the decomposition, stage shapes and exchange
geometry the ParaFaFT drivers compute, restated for this package.

decompose               parafaft_common.hpp:73-80 (Barry Smith's balanced split)
r2c_stage_shapes        parafaft_r2c.hpp:1069-1151
exchange_geometry       parafaft_common.hpp:345-393 (ExchangeGeometry, :273-304)
"""
from __future__ import annotations

import math
from dataclasses import dataclass


def decompose(N: int, M: int, p: int):
    """(n, s): part p of N items over M parts; the first N % M parts hold one
    extra item."""
    q, r = divmod(N, M)
    n = q + 1 if r > p else q
    s = n * p if r > p else q * p + r
    return n, s


def r2c_stage_shapes(global_real, dims, coords):
    """Per-stage local shapes of a ParaFaFT_R2C rank at `coords` on the grid
    `dims` (D-1 entries).  Returns (input_shapes, output_shapes, real_start,
    complex_start); stage 0 input is the real shape, every other entry is a
    complex shape with the last global axis n//2+1."""
    D = len(global_real)
    gc = list(global_real)
    gc[-1] = global_real[-1] // 2 + 1
    local = [decompose(global_real[i], dims[i], coords[i]) for i in range(D - 1)]
    ins = [[0] * D for _ in range(D)]
    outs = [[0] * D for _ in range(D)]
    for a in range(D - 1):
        ins[0][a] = outs[0][a] = local[a][0]
    ins[0][D - 1] = global_real[D - 1]
    outs[0][D - 1] = gc[D - 1]
    for stage in range(1, D):
        for axis in range(D):
            if axis < D - 1 - stage:
                v = local[axis][0]
            elif axis == D - 1 - stage:
                v = gc[axis]
            else:
                grid_dim = (D - 1 - stage) + (axis - (D - stage))
                if 0 <= grid_dim < D - 1:
                    v = decompose(gc[axis], dims[grid_dim], coords[grid_dim])[0]
                else:
                    v = gc[axis]
            ins[stage][axis] = outs[stage][axis] = v
    real_start = [local[i][1] for i in range(D - 1)] + [0]
    complex_start = [0] + [decompose(gc[i], dims[i - 1], coords[i - 1])[1] for i in range(1, D)]
    return ins, outs, real_start, complex_start


@dataclass
class ExchangeGeometry:
    src_n: list
    src_s: list
    dst_n: list
    dst_s: list
    send_counts: list
    send_displs: list
    recv_counts: list
    recv_displs: list
    src_leading: int
    src_trailing: int
    dst_leading: int
    dst_trailing: int
    src_axis_extent: int
    dst_axis_extent: int

    @property
    def src_contiguous(self):
        return self.src_leading == 1

    @property
    def dst_contiguous(self):
        return self.dst_leading == 1


def exchange_geometry(nparts, src_sizes, src_axis, dst_sizes, dst_axis) -> ExchangeGeometry:
    nd = len(src_sizes)
    sl = math.prod(src_sizes[:src_axis])
    st = math.prod(src_sizes[src_axis + 1:nd])
    dl = math.prod(dst_sizes[:dst_axis])
    dt = math.prod(dst_sizes[dst_axis + 1:nd])
    src = [decompose(src_sizes[src_axis], nparts, p) for p in range(nparts)]
    dst = [decompose(dst_sizes[dst_axis], nparts, p) for p in range(nparts)]
    sc = [sl * n * st for n, _ in src]
    rc = [dl * n * dt for n, _ in dst]
    sd = [sum(sc[:p]) for p in range(nparts)]
    rd = [sum(rc[:p]) for p in range(nparts)]
    return ExchangeGeometry([n for n, _ in src], [s for _, s in src], [n for n, _ in dst],
                            [s for _, s in dst], sc, sd, rc, rd, sl, st, dl, dt,
                            src_sizes[src_axis], dst_sizes[dst_axis])
