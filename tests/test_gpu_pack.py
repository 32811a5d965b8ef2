"""Bit-exact pack/unpack against the oracle's restatement of the reference's
loops (parafaft_common.hpp:576-590, 654-668) and against the independent
per-element C-order subarray enumerator (parafaft_common.hpp:108-160)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu


def _geom_arrays(g, pack):
    n = g.src_n if pack else g.dst_n
    s = g.src_s if pack else g.dst_s
    d = g.send_displs if pack else g.recv_displs
    return ((C.c_int * len(n))(*n), (C.c_int * len(s))(*s), (C.c_uint64 * len(d))(*d))


def _dev_pack(src_dev, g, esz, total_elems):
    import torch
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    out = torch.zeros(total_elems * esz, dtype=torch.uint8, device="cuda")
    n, s, d = _geom_arrays(g, True)
    A.check(A.load().pfx_pack(out.data_ptr(), src_dev.data_ptr(), len(g.src_n), n, s, d,
                              g.src_leading, g.src_trailing, g.src_axis_extent, esz, None),
            "pfx_pack")
    return out


def _dev_unpack(recv_dev, dst_dev, g, esz):
    from parafaft_amd import _capi as A
    n, s, d = _geom_arrays(g, False)
    A.check(A.load().pfx_unpack(dst_dev.data_ptr(), recv_dev.data_ptr(), len(g.dst_n), n, s, d,
                                g.dst_leading, g.dst_trailing, g.dst_axis_extent, esz, None),
            "pfx_unpack")


CASES = [
    # (src shape, src axis, dst shape, dst axis, nparts, esz)
    ([8, 16, 33], 2, [8, 64, 9], 1, 4, 8),      # uneven split 33 -> 9,8,8,8
    ([6, 10, 12], 1, [24, 10, 3], 0, 4, 16),
    ([5, 7, 11, 13], 3, [5, 7, 44, 4], 2, 4, 8),
    ([16, 16, 1025 // 8], 2, [16, 64, 33], 1, 4, 8),
    ([3, 4], 1, [12, 1], 0, 1, 8),              # nparts == 1
    ([9, 9, 9], 2, [9, 18, 5], 1, 2, 4),        # 4-byte elements
]


@pytest.mark.parametrize("case", CASES)
def test_pack_unpack_bit_exact(case):
    import torch
    src_shape, sa, dst_shape, da, nparts, esz = case
    g = O.exchange_geometry(nparts, src_shape, sa, dst_shape, da)
    nsrc = int(np.prod(src_shape))
    rng = np.random.default_rng(len(src_shape) * 100 + nparts)
    src = rng.integers(0, 256, nsrc * esz, dtype=np.uint8)
    ref = O.pack(src, g, esz)
    # independent check of the oracle: concatenated subarray serialisations
    ser = np.concatenate([O.subarray_serialize(src, src_shape, sa, nparts, p, esz)
                          for p in range(nparts)])
    assert np.array_equal(ref, ser)
    dsrc = to_dev(src)
    got = to_host(_dev_pack(dsrc, g, esz, sum(g.send_counts)))
    assert np.array_equal(got, ref)
    # unpack into a destination pre-filled with a sentinel
    ndst = int(np.prod(dst_shape))
    recv = rng.integers(0, 256, sum(g.recv_counts) * esz, dtype=np.uint8)
    dst0 = np.full(ndst * esz, 0xA5, dtype=np.uint8)
    refd = O.unpack(recv, dst0.copy(), g, esz)
    ddst = to_dev(dst0)
    _dev_unpack(to_dev(recv), ddst, g, esz)
    assert np.array_equal(to_host(ddst), refd)


def test_config3_transition_geometry_small():
    """The config-3 transitions (2048^3 R2C on a 4x4 grid, rank (0,0)) scaled
    down by 8 on the non-split axes: T0 [64,64,129]/axis2 -> [64,256,33]/axis1,
    T1 [64,256,33]/axis1 -> [256,64,33]/axis0, c64 elements."""
    for (ss, sa, ds, da) in [([64, 64, 129], 2, [64, 256, 33], 1),
                             ([64, 256, 33], 1, [256, 64, 33], 0)]:
        g = O.exchange_geometry(4, ss, sa, ds, da)
        nsrc = int(np.prod(ss))
        src = np.random.default_rng(3).integers(0, 256, nsrc * 8, dtype=np.uint8)
        ref = O.pack(src, g, 8)
        got = to_host(_dev_pack(to_dev(src), g, 8, sum(g.send_counts)))
        assert np.array_equal(got, ref)
