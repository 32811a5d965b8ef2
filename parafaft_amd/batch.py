"""Many independent transforms of uneven shapes in grouped launches, and HIP
graph capture (pfx_batch_* / pfx_graph_* of the C ABI)."""
from __future__ import annotations

import ctypes as C

from . import _capi as A
from .backend import precision_code, ptr

KIND = {"c2c": A.PFX_C2C, "r2c": A.PFX_R2C}


class Batch:
    """`items`: iterable of (kind, shape, device pointer); kind "c2c" or "r2c"
    (padded real array, last axis 2*(n/2+1) scalars)."""

    def __init__(self, precision, items):
        items = list(items)
        arr = (A.BatchDesc * max(1, len(items)))()
        for i, (kind, shape, data) in enumerate(items):
            arr[i].kind = KIND[kind] if isinstance(kind, str) else kind
            arr[i].ndim = len(shape)
            for k, v in enumerate(shape):
                arr[i].shape[k] = int(v)
            arr[i].data = ptr(data)
        h = C.c_void_p()
        A.check(A.load().pfx_batch_create(C.byref(h), precision_code(precision), len(items),
                                          C.cast(arr, C.c_void_p)), "pfx_batch_create")
        self.handle = h.value
        self.count = len(items)

    def tune(self, flag=A.PFX_PLAN_MEASURE, stream=None):
        """Plan every transform with `flag` on its own buffer (contents are
        lost) and regroup the launches around the chosen kernels."""
        A.check(A.load().pfx_batch_tune(self.handle, flag, ptr(stream)), "pfx_batch_tune")

    def execute(self, direction, stream=None):
        A.check(A.load().pfx_batch_execute(self.handle, direction, ptr(stream)),
                "pfx_batch_execute")

    def execute_each(self, direction, stream=None):
        A.check(A.load().pfx_batch_execute_each(self.handle, direction, ptr(stream)),
                "pfx_batch_execute_each")

    def execute_persistent(self, direction, stream=None):
        """All transforms in one persistent launch (tickets from per-XCD
        queues, device-side stage dependencies)."""
        A.check(A.load().pfx_batch_execute_persistent(self.handle, direction, ptr(stream)),
                "pfx_batch_execute_persistent")

    def persistent_failed(self, stream=None) -> bool:
        """Synchronises `stream`; True if the last persistent launch stopped
        at its spin limit."""
        f = C.c_int()
        A.check(A.load().pfx_batch_persistent_status(self.handle, ptr(stream), C.byref(f)),
                "pfx_batch_persistent_status")
        return bool(f.value)

    def set_spin_limit(self, polls: int):
        A.check(A.load().pfx_batch_set_spin_limit(self.handle, int(polls)),
                "pfx_batch_set_spin_limit")

    def persistent_info(self, direction):
        """(jobs, tickets, workgroups) of the persistent schedule."""
        j, t, w = C.c_int(), C.c_int(), C.c_int()
        A.check(A.load().pfx_batch_persistent_info(self.handle, direction, C.byref(j),
                                                   C.byref(t), C.byref(w)),
                "pfx_batch_persistent_info")
        return j.value, t.value, w.value

    def launches(self, direction):
        g, e = C.c_int(), C.c_int()
        A.check(A.load().pfx_batch_launches(self.handle, direction, C.byref(g), C.byref(e)),
                "pfx_batch_launches")
        return g.value, e.value

    def __del__(self):
        try:
            if self.handle:
                A.load().pfx_batch_destroy(self.handle)
                self.handle = 0
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph of everything `fn()` enqueues on `stream`."""

    def __init__(self, stream, fn):
        L = A.load()
        s = ptr(stream)
        A.check(L.pfx_graph_begin_capture(s), "pfx_graph_begin_capture")
        try:
            fn()
        finally:
            h = C.c_void_p()
            A.check(L.pfx_graph_end_capture(s, C.byref(h)), "pfx_graph_end_capture")
        self.handle = h.value

    def launch(self, stream):
        A.check(A.load().pfx_graph_launch(self.handle, ptr(stream)), "pfx_graph_launch")

    def __del__(self):
        try:
            if self.handle:
                A.load().pfx_graph_destroy(self.handle)
                self.handle = 0
        except Exception:
            pass


def uneven_workload(seed: int = 20241203, count: int = 512):
    """The uneven batch: 40% 3D C2C cubes {48,64,96,128}^3, 30% 2D C2C
    {1024,2048,4096}^2, 20% 3D R2C {80,120,160}^3, 10% 4D C2C 32^4, shapes
    drawn from `seed`.  Returns a list of (kind, shape)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n40, n30, n20 = round(0.4 * count), round(0.3 * count), round(0.2 * count)
    n10 = count - n40 - n30 - n20
    items = ([("c2c", (int(n),) * 3) for n in rng.choice([48, 64, 96, 128], n40)] +
             [("c2c", (int(n),) * 2) for n in rng.choice([1024, 2048, 4096], n30)] +
             [("r2c", (int(n),) * 3) for n in rng.choice([80, 120, 160], n20)] +
             [("c2c", (32,) * 4)] * n10)
    order = rng.permutation(len(items))
    return [items[i] for i in order]


def storage_elems(kind: str, shape) -> int:
    """Complex elements of in-place storage (R2C: padded last axis)."""
    n = 1
    for v in shape[:-1]:
        n *= v
    last = shape[-1] // 2 + 1 if kind == "r2c" else shape[-1]
    return n * last
