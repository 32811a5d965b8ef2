"""CPU checks of the C ABI: libpfx.so loads, exports every symbol that
include/pfx_capi.h declares, and validates arguments on the host (no device
work is launched here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(pfx_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    import parafaft_amd as P
    from parafaft_amd import _capi as A
    L = P.load_library()
    declared = _declared_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    bound = {n for n, _, _ in A.SIGNATURES}
    assert declared == bound, declared ^ bound


def test_error_strings_and_version():
    import parafaft_amd as P
    L = P.load_library()
    assert L.pfx_error_string(0) == b"success"
    assert L.pfx_error_string(2) == b"unsupported configuration"
    assert b"gfx950" in L.pfx_version()


@pytest.mark.parametrize("args,code", [
    ((0, 0, 4, 1, 0), 1),        # length 0
    ((0, 16, 4, 0, 16), 1),      # stride 0
    ((7, 16, 4, 1, 16), 2),      # unknown precision
    ((0, 16, -1, 1, 16), 1),     # negative batch
])
def test_plan_argument_validation(args, code):
    import parafaft_amd as P
    L = P.load_library()
    h = C.c_void_p()
    prec, n, batch, stride, dist = args
    rc = L.pfx_plan_create_c2c(C.byref(h), prec, n, batch, stride, dist)
    assert rc == code
    assert h.value is None
    assert L.pfx_last_error_message()


def test_r2c_validation():
    import parafaft_amd as P
    L = P.load_library()
    h = C.c_void_p()
    assert L.pfx_plan_create_r2c(C.byref(h), 0, 16, 4, 2, 18) == 2      # stride != 1
    assert L.pfx_plan_create_r2c(C.byref(h), 0, 16, 4, 1, 16) == 1      # dist < 2(n/2+1)
    assert L.pfx_plan_create_c2r(C.byref(h), 0, 16, 4, 1, 17) == 1      # odd dist


def test_python_mirror_raises_pfx_error():
    import parafaft_amd as P
    with pytest.raises(P.PfxError) as e:
        P.Plan.c2c("f32", 0, 1, 1, 1)
    assert "length" in str(e.value)


def test_execute_and_pack_reject_bad_input_without_device_work():
    import parafaft_amd as P
    L = P.load_library()
    assert L.pfx_execute(None, 0, None, None) == 1
    n = (C.c_int * 2)(3, 4)
    s = (C.c_int * 2)(0, 3)
    d = (C.c_uint64 * 2)(0, 3)
    # block [3, 7) exceeds an axis extent of 6
    assert L.pfx_pack(None, None, 2, n, s, d, 1, 1, 6, 8, None) == 1
    assert L.pfx_scale(None, 5, 0, 2.0, None) == 1
    assert L.pfx_ndplan_create(C.byref(C.c_void_p()), 0, 1, 0, None) == 1


def test_no_cpu_fallback_in_product():
    """The shipped package never imports the oracle."""
    pkg = os.path.join(ROOT, "parafaft_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.replace("ORACLE", ""), f


def test_batch_product_overflow_rejected():
    """A two-level batch whose transform count overflows int64 is refused on
    the host, before any allocation or device work."""
    import parafaft_amd as P
    L = P.load_library()
    h = C.c_void_p()
    rc = L.pfx_plan_create_c2c_many(C.byref(h), 0, 16, 1 << 40, 1, 16, 1 << 40, 1 << 44)
    assert rc == 1
    assert h.value is None
    assert b"overflows" in L.pfx_last_error_message()


@pytest.mark.parametrize("prec,n", [(0, (1 << 23) + 1), (1, (1 << 23) + 9), (0, 16777259)])
def test_lengths_without_a_kernel_are_refused_on_the_host(prec, n):
    """C2C lengths with no tiled plan beyond the direct-DFT limit that
    neither split into two tiled lengths nor fit the chirp convolution
    (n <= 2^23) are refused before any device work, with
    PFX_ERR_UNSUPPORTED; the GPU tests hold the complementary rule (every
    other length must plan)."""
    import sys
    sys.path.insert(0, ROOT)
    from tests.gpu_util import c2c_supported
    import parafaft_amd as P
    L = P.load_library()
    assert not c2c_supported("f64" if prec else "f32", n)
    h = C.c_void_p()
    assert L.pfx_plan_create_c2c(C.byref(h), prec, n, 4, 1, n) == 2
    assert h.value is None
    assert b"fallback" in L.pfx_last_error_message()
    assert b"no split into two tiled lengths" in L.pfx_last_error_message()
    assert b"chirp" in L.pfx_last_error_message()


def test_long_split_rule():
    """The four-step split the tests expect: the most even pair of tiled
    factors (the library's rule in make_long)."""
    import sys
    sys.path.insert(0, ROOT)
    from tests.gpu_util import c2c_supported, long_split, real_supported
    assert long_split(8192) == (64, 128)
    assert long_split(16384) == (128, 128)
    assert long_split(6144) == (64, 96)
    assert long_split(8191) is None and long_split(4097) is None
    assert c2c_supported("f64", 3072) and c2c_supported("f32", 1 << 20)
    assert c2c_supported("f32", 8191) and not c2c_supported("f32", (1 << 23) + 1)
    assert real_supported("f32", 32768) and real_supported("f32", 8194)
    assert real_supported("f64", 4097) and not real_supported("f32", 2 * ((1 << 23) + 1))


def test_chirp_length_rule():
    """The chirp plans' convolution length the tests expect (the library's
    rule in chirp_length): one-kernel lengths of the forms b 2^k first,
    then the shortest that splits."""
    import sys
    sys.path.insert(0, ROOT)
    from tests.gpu_util import chirp_length
    assert chirp_length("f32", 1100) == 2560      # 5 * 2^9 instead of 4096
    assert chirp_length("f32", 1001) == 2048
    assert chirp_length("f32", 700) == 1536
    assert chirp_length("f32", 4099) == 16384     # one kernel beats a four-step 12288
    assert chirp_length("f64", 4099) == 16384
    assert chirp_length("f32", 10007) == 20480    # no one-kernel length: shortest split
    assert chirp_length("f64", 9000) == 20480
    assert chirp_length("f32", (1 << 23)) == 1 << 24


def test_tune_cache_load_counts_dropped_lines(tmp_path):
    """Malformed keys and kernels this build lacks are dropped and counted
    (host-only: no plan is created)."""
    from parafaft_amd import backend as B
    path = tmp_path / "cache.txt"
    path.write_text("# header\n"
                    "garbage\tvalue\n"
                    "p0 k0 n16 i4 o1 s1 d16 do0\t0|0|16|0|1|nope|order=1\n"
                    "no tab\n")
    B.tune_cache_clear()
    assert B.tune_cache_load(str(path)) == 3
    assert B.tune_cache_size() == 0
    import parafaft_amd as P
    assert b"3 line(s) dropped" in P.load_library().pfx_last_error_message()


def test_tune_cache_keeps_in_context_chunk_choices(tmp_path):
    """An N-D plan's in-context chunk choice is cached under the chunk
    plan's key plus " chunkP ofN" (forward) or " chunkP ofN bwd" (backward
    chunk plans): such lines load (host-only); a truncated or unknown suffix
    is dropped."""
    from parafaft_amd import backend as B
    path = tmp_path / "cache.txt"
    val = "0|0|1024|1|32|wide-nt:32, 32|order=32"
    key = "p0 k0 n1024 i1024 o32 s1024 d1 do1048576"
    path.write_text(f"{key} chunk32 of1024\t{val}\n"
                    f"{key} chunk32\t{val}\n"
                    f"{key} chunk32 of1024 bwd\t{val}\n"
                    f"{key} chunk32 of1024 fwd\t{val}\n"
                    f"{key} chunk32 of1024 bwdx\t{val}\n")
    B.tune_cache_clear()
    assert B.tune_cache_load(str(path)) == 3
    assert B.tune_cache_size() == 2
    B.tune_cache_clear()
