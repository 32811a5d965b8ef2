"""pfx_execute_segmented on plain C2C plans with hand-made segment maps: up
to 16 peers (uneven and empty blocks), transform lengths up to 4096, row and
batch-contiguous layouts, loads from segments, stores into segments, or
both, for every candidate kernel of the plan.  The result must equal, bit
for bit, the same kernel's plain transform followed by the pack of
parafaft_common.hpp:576-590 (peer p's block of the split axis, C order), or
the unpack followed by the plain transform."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu

CASES = [  # (n, lead, trailing, counts)
    (4096, 3, 1, [4096]),                                    # one peer: offsets up to 4095
    (4096, 2, 1, [300, 0, 1000, 7, 0, 1, 2000, 788] + [0] * 8),  # 16 peers, empty blocks
    (1024, 3, 33, [64] * 16),                                # 16 equal peers, batch-contiguous
    (2048, 2, 17, [513, 512, 512, 511]),                     # decompose-style split
    (2048, 1, 257, [1, 2046, 1]),                            # one-element blocks
    (4096, 2, 9, [1024, 0, 3072]),                           # kernels without table room
]


def _maps(counts, segs):
    from parafaft_amd import _capi as A
    k = len(counts)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
    arrays = ((C.c_void_p * k)(*segs), (C.c_int * k)(*starts.tolist()),
              (C.c_int * k)(*[int(c) for c in counts]))
    return arrays, starts


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("side", ["out", "in", "both"])
def test_segmented_bit_exact(prec, case, side):
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    n, lead, trailing, counts = CASES[case]
    assert sum(counts) == n
    cdt = np.complex128 if prec == "f64" else np.complex64
    x = O.uniform_complex((lead, n, trailing), seed=case * 7 + len(side)).astype(cdt)
    plan = P.Plan.c2c(prec, n, trailing, trailing, 1, lead, n * trailing)
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])

    def blocks(a):   # per-peer blocks of the split axis, C order
        return [np.ascontiguousarray(a[:, s:s + c, :]).reshape(-1) for s, c in zip(starts, counts)]

    blocks_x = blocks(x)
    # device segments (at least one element each, so every peer has an address)
    seg_in = [to_dev(b if b.size else np.zeros(1, cdt)) for b in blocks_x]
    seg_out = [to_dev(np.zeros(max(b.size, 1), cdt)) for b in blocks_x]
    data = to_dev(x.reshape(-1) if side == "out" else np.zeros(x.size, cdt))
    arr_in, _ = _maps(counts, [t.data_ptr() for t in seg_in])
    arr_out, _ = _maps(counts, [t.data_ptr() for t in seg_out])
    m_in = A.SegMap(len(counts), arr_in[0], arr_in[1], arr_in[2], trailing)
    m_out = A.SegMap(len(counts), arr_out[0], arr_out[1], arr_out[2], trailing)
    use_in = side in ("in", "both")
    use_out = side in ("out", "both")
    ran = 0
    for i in range(plan.num_candidates()):
        plan.use_candidate(i)
        name = plan.describe()[0]
        # reference: this kernel's plain transform, then the pack
        d = to_dev(x.reshape(-1))
        plan.execute(A.PFX_FORWARD, d)
        y = to_host(d).reshape(lead, n, trailing)
        blocks_y = blocks(y)
        for t in seg_out:
            t.zero_()
        data.copy_(to_dev(x.reshape(-1) if side == "out" else np.zeros(x.size, cdt)))
        rc = A.load().pfx_execute_segmented(
            plan.handle, A.PFX_FORWARD, data.data_ptr(),
            C.byref(m_in) if use_in else None, C.byref(m_out) if use_out else None, 1.0, None)
        if rc == A.PFX_ERR_UNSUPPORTED:   # a candidate with no segmented form
            continue
        A.check(rc, "pfx_execute_segmented")
        ran += 1
        if use_out:
            for p, want in enumerate(blocks_y):
                got = to_host(seg_out[p])[:want.size]
                assert np.array_equal(got, want), (name, "segment", p)
        else:
            assert np.array_equal(to_host(data).reshape(y.shape), y), name
    assert ran >= 1
