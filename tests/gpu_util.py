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
