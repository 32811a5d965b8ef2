"""Helpers shared by the GPU parity tests: device buffers via torch (plumbing
only), the library through the C ABI, the oracle as the checker."""
import numpy as np


def torch():
    import torch as t
    return t


def to_dev(a: np.ndarray):
    t = torch()
    x = t.from_numpy(np.ascontiguousarray(a)).to("cuda")
    t.cuda.synchronize()
    return x


def to_host(x) -> np.ndarray:
    t = torch()
    t.cuda.synchronize()
    return x.cpu().numpy()


def rel_l2(y, ref) -> float:
    y = np.asarray(y).astype(np.complex128).ravel()
    ref = np.asarray(ref).astype(np.complex128).ravel()
    den = np.sqrt(np.sum(np.abs(ref) ** 2))
    return float(np.sqrt(np.sum(np.abs(y - ref) ** 2)) / (den if den > 0 else 1.0))


def tol(prec: str, n: int) -> float:
    """FP64: rel-L2 <= 1e-13; FP32: rel-L2 <= 1e-6 * log2(N) (N >= 2)."""
    if prec == "f64":
        return 1e-13
    return 1e-6 * max(1.0, np.log2(max(n, 2)))


# Which lengths the library must accept: every product of a radix plan in
# pfx_plans.inc has tiled kernels (C2C, and R2C/C2R of twice the length), any
# length up to the direct-DFT limit (2n complex values in 64 KiB of LDS) runs
# the fallback, a longer length that splits into two tiled lengths runs the
# four-step plan, and any other length up to 2^23 Bluestein's chirp
# convolution (an even real length: its half as C2C; an odd one: the whole
# row as C2C).  A plan outside these must be refused; every other plan must
# be created (a refusal there is a failure, not a skip).
NAIVE_MAX = {"f32": 4096, "f64": 2048}


def tiled_lengths():
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "parafaft_amd", "csrc", "pfx_plans.inc")
    out = set()
    for m in re.finditer(r"^PFX_PLAN\(\s*\d+\s*,([\d,\s]+)\)", open(path).read(), re.M):
        n = 1
        for r in m.group(1).split(","):
            n *= int(r)
        out.add(n)
    return out


def long_split(n: int):
    """The four-step split of a long length: the most even n1 * n2 = n with
    both factors tiled (n1 <= n2), or None."""
    t = tiled_lengths()
    best = None
    f = 2
    while f * f <= n:
        if n % f == 0 and f in t and n // f in t:
            best = (f, n // f)
        f += 1
    return best


# Row lengths beyond the tiled plans that still run one kernel.
LONG_ROW_ONE_KERNEL = {"f32": {8192, 16384}, "f64": {8192, 16384}}


def _split_ok(m: int) -> bool:
    t = tiled_lengths()
    f = 2
    while f * f <= m:
        if m % f == 0 and f in t and m // f in t:
            return True
        f += 1
    return False


def chirp_length(prec: str, n: int) -> int:
    """Convolution length of a chirp plan: the shortest M >= 2n - 1 of the
    forms b 2^k (b = 1, 3, 5, 15, 25, 125) with a one-kernel row plan, else
    the shortest such M that splits into two tiled lengths."""
    base = "f64" if prec == "f64" else "f32"
    one = tiled_lengths() | LONG_ROW_ONE_KERNEL[base]
    need = 2 * n - 1
    cands = []
    for b in (1, 3, 5, 15, 25, 125):
        m = b
        while m < need:
            m *= 2
        cands.append(m)
    ones = [m for m in cands if m in one]
    if ones:
        return min(ones)
    return min(m for m in cands if _split_ok(m))


CHIRP_MAX = 1 << 23   # Bluestein: the M >= 2n - 1 convolution must split (M <= 4096^2)


def c2c_supported(prec: str, n: int) -> bool:
    return n in tiled_lengths() or n <= NAIVE_MAX["f64" if prec == "f64" else "f32"] or \
        long_split(n) is not None or n <= CHIRP_MAX


def real_supported(prec: str, n: int) -> bool:
    return (n % 2 == 0 and n // 2 in tiled_lengths()) or \
        n <= NAIVE_MAX["f64" if prec == "f64" else "f32"] or \
        (n % 2 == 0 and c2c_supported(prec, n // 2)) or (n % 2 == 1 and c2c_supported(prec, n))


def forced_sub_plan_kernels(prec, n, batch, stride, dist, cache_path):
    """Every kernel candidate of every tiled sub-plan of a long plan, forced
    in turn through a tuning-cache line: yields (sub-plan key, kernel name,
    long plan created with that line loaded).  The long plan is first
    Measure-tuned so the cache holds one line per sub-plan geometry."""
    import parafaft_amd as P
    from parafaft_amd import backend as B
    B.tune_cache_clear()
    plan = P.Plan.c2c(prec, n, batch, stride, dist)
    size = (batch - 1) * dist + (n - 1) * stride + 1
    esz = np.complex128 if prec == "f64" else np.complex64
    plan.tune(1, to_dev(np.zeros(size, esz)))
    B.tune_cache_save(cache_path)
    lines = [l.split("\t") for l in open(cache_path).read().splitlines()
             if l and not l.startswith("#")]
    # one per tiled sub-plan (a batch-contiguous four-step has one: its
    # first pass is a fused kernel of its own)
    assert len(lines) >= 1, lines
    for key, val in lines:
        # key: "p<prec> k<kind> n<length> i<inner> o<outer> s<stride> d<dist> do<dist_outer>"
        f = {}
        for tok in key.split():
            name = "do" if tok.startswith("do") else tok[0]
            f[name] = int(tok[len(name):])
        sub = P.Plan.c2c(prec, f["n"], f["i"], f["s"], f["d"], f["o"], f["do"])
        prefix = "|".join(val.split("|")[:4])  # precision, kind, length, layout
        for i in range(sub.num_candidates()):
            sub.use_candidate(i)
            name = sub.describe()[0]
            rad = name[name.index("_r[") + 3:name.index("]_c")]
            per_block = int(name[name.index("]_c") + 3:].split("_")[0])
            kid = f"{prefix}|{per_block}|{rad}|order=1"
            with open(cache_path, "w") as fh:
                for k2, v2 in lines:
                    fh.write(f"{k2}\t{kid if k2 == key else v2}\n")
            B.tune_cache_clear()
            assert B.tune_cache_load(cache_path) == 0, (key, kid)
            yield key, name, P.Plan.c2c(prec, n, batch, stride, dist)
    B.tune_cache_clear()
