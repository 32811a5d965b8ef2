"""Pins the oracle's geometry restatements (CPU only) with the reference's own
unit-test known answers (test/unit/test_exchange_geometry.cpp), the worked
examples in the reference's comments, brute-force enumeration of the MPI
C-order subarray, and the invariants of the Barry Smith decomposition."""
import numpy as np
import pytest

from oracle import oracle as O

INT_MAX = 2 ** 31 - 1


@pytest.mark.parametrize("N,M", [(1025, 4), (2048, 4), (7, 3), (10, 10), (5, 8), (1, 1), (31, 7)])
def test_decompose_invariants(N, M):
    parts = [O.decompose(N, M, p) for p in range(M)]
    assert sum(n for n, _ in parts) == N
    pos = 0
    for n, s in parts:
        assert s == pos and n in (N // M, N // M + 1)
        pos += n
    # the first N % M parts take the extra element (parafaft_common.hpp:73-80)
    assert [n for n, _ in parts] == [N // M + (1 if p < N % M else 0) for p in range(M)]


def test_decompose_config3_known_answer():
    assert [O.decompose(1025, 4, p) for p in range(4)] == [(257, 0), (256, 257), (256, 513),
                                                          (256, 769)]


@pytest.mark.parametrize("total,nt", [(10, 3), (3, 8), (0, 4), (1024, 16), (7, 7)])
def test_chunk_range_partitions(total, nt):
    rng = [O.chunk_range(total, t, nt) for t in range(nt)]
    assert rng[0][0] == 0 and rng[-1][1] == total
    for (b0, e0), (b1, _) in zip(rng, rng[1:]):
        assert e0 == b1
    sizes = [e - b for b, e in rng]
    assert max(sizes) - min(sizes) <= 1


def test_stage_shapes_doc_example():
    """parafaft_c2c.hpp:774-778: [32,32,32] on a [2,2] grid."""
    shapes, start = O.stage_shapes_c2c([32, 32, 32], [2, 2], [0, 0])
    assert shapes == [[16, 16, 32], [16, 32, 16], [32, 16, 16]]
    assert start == [0, 0, 0]
    shapes, start = O.stage_shapes_c2c([32, 32, 32], [2, 2], [1, 1])
    assert start == [16, 16, 0]


def test_stage_shapes_one_rank_are_global():
    for shp in ([64, 64, 64], [16, 16, 16, 16], [512, 512], [8, 8, 8, 8, 8]):
        shapes, _ = O.stage_shapes_c2c(shp, [1] * (len(shp) - 1), [0] * (len(shp) - 1))
        assert all(s == shp for s in shapes)


def test_stage_shapes_noncubic_reference_drift():
    """ParaFaFT_C2C::setup_stage_shapes sizes a redistributed axis with
    local_n[grid_dim] (parafaft_c2c.hpp:818-825), i.e. the split extent of a
    different axis.  That equals decompose(global_shape[axis], ...) only when
    the extents agree (every reference test is a hypercube); on a non-cubic
    grid the reference's own stage shape is wrong.  The oracle keeps the
    reference formula; the one-rank drivers of this build use the true shape,
    which on one rank is the global shape."""
    shapes, _ = O.stage_shapes_c2c([12, 20, 10], [1, 1], [0, 0])
    assert shapes[0] == [12, 20, 10]
    assert shapes[1] == [12, 20, 20]       # reference formula: local_n[1] = 20, not 10
    assert shapes[2] == [12, 12, 20]


def test_r2c_stage_shapes_config3():
    """2048^3 R2C on a 4x4 grid at coords (0,0): the pencil share the
    benchmark's config 3 runs on one GPU."""
    ins, outs, rs, cs = O.stage_shapes_r2c([2048, 2048, 2048], [4, 4], [0, 0])
    assert ins[0] == [512, 512, 2048] and outs[0] == [512, 512, 1025]
    assert ins[1] == outs[1] == [512, 2048, 257]
    assert ins[2] == outs[2] == [2048, 512, 257]
    assert rs == [0, 0, 0] and cs == [0, 0, 0]
    ins, outs, rs, cs = O.stage_shapes_r2c([2048, 2048, 2048], [4, 4], [3, 1])
    assert ins[2] == [2048, 512, 256] and cs == [0, 1536, 257] and rs == [1536, 512, 0]


def test_exchange_geometry_large_displacements():
    """test_exchange_geometry.cpp:53-84: 2048^3 split in 2 along axis 2 ->
    per-part count 2^32, second displacement 2^32 (no int wrap)."""
    g = O.exchange_geometry(2, [2048, 2048, 2048], 2, [2048, 2048, 2048], 1)
    expected = 2048 * 2048 * 1024
    assert g.send_counts[0] == expected > INT_MAX
    assert g.send_displs == [0, expected]
    assert g.recv_displs[1] > INT_MAX


def test_exchange_geometry_identity_at_one_part():
    """test_exchange_geometry.cpp:89-108."""
    g = O.exchange_geometry(1, [8, 16, 32], 2, [8, 16, 32], 1)
    assert g.src_n[0] == g.src_axis_extent and g.dst_n[0] == g.dst_axis_extent
    assert g.src_n[0] * g.src_trailing * g.src_leading == g.dst_n[0] * g.dst_trailing * g.dst_leading


def test_exchange_geometry_config3_t0():
    g = O.exchange_geometry(4, [512, 512, 1025], 2, [512, 2048, 257], 1)
    assert g.src_n == [257, 256, 256, 256]
    assert g.send_counts == [512 * 512 * 257] + [512 * 512 * 256] * 3
    assert g.recv_counts == [512 * 512 * 257] * 4
    assert g.src_leading == 512 * 512 and g.src_trailing == 1 and g.src_axis_extent == 1025
    assert g.dst_leading == 512 and g.dst_trailing == 257


@pytest.mark.parametrize("seed", range(12))
def test_pack_equals_subarray_serialisation(seed):
    """The pack loops produce exactly MPI's C-order subarray wire order
    (parafaft_common.hpp:148-158 vs 576-590), for random shapes and splits."""
    rng = np.random.default_rng(seed)
    nd = int(rng.integers(2, 5))
    sizes = [int(v) for v in rng.integers(1, 9, nd)]
    axis = int(rng.integers(0, nd))
    nparts = int(rng.integers(1, sizes[axis] + 1))
    esz = int(rng.choice([1, 4, 8, 16]))
    src = rng.integers(0, 256, int(np.prod(sizes)) * esz, dtype=np.uint8)
    g = O.exchange_geometry(nparts, sizes, axis, sizes, axis)
    packed = O.pack(src, g, esz)
    ser = np.concatenate([O.subarray_serialize(src, sizes, axis, nparts, p, esz)
                          for p in range(nparts)])
    assert np.array_equal(packed, ser)
    # unpack is the inverse of pack on the same geometry
    back = O.unpack(packed, np.zeros_like(src), g, esz)
    assert np.array_equal(back, src)


def test_generators_known_values():
    """test_helpers.hpp:106-221."""
    g = O.generate_shape(3, 8, 0)
    assert g[4, 4, 4] == 1.0 and abs(g[0, 4, 4] - np.exp(-16 / 32)) < 1e-15
    s = O.generate_shape(2, 8, 1)
    assert s[4, 4] == 1.0 and s[4, 6] == 1.0 and s[4, 7] == 0.0 and s[0, 0] == 0.0
    # LCG: seed = 42 + 7 D + 13 N, first coefficient
    D, N = 2, 8
    seed = (42 + 7 * D + 13 * N) & 0xFFFFFFFF
    coeffs = []
    for _ in range(1 + 3 * D):
        seed = (seed * 1103515245 + 12345) & 0xFFFFFFFF
        coeffs.append(((seed >> 16) & 0x7FFF) / 32768.0 - 0.5)
    p = O.generate_shape(D, N, 2)
    x0, x1 = 3 / N, 5 / N
    v = coeffs[0] + coeffs[1] * x0 + coeffs[2] * x1 + coeffs[3] * x0 ** 2 + coeffs[4] * x1 ** 2 \
        + coeffs[5] * x0 ** 3 + coeffs[6] * x1 ** 3
    assert abs(p[3, 5] - v) < 1e-15


def test_uniform_generator_range_and_determinism():
    a = O.uniform(10000, 1)
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.05
    assert np.array_equal(a, O.uniform(10000, 1))
    assert not np.array_equal(a, O.uniform(10000, 2))
