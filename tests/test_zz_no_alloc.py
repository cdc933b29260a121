"""nbhip_accelerate_noalloc: the direct sum over the caller's device arrays, with its
scratch in a static device buffer (no allocation), through the C ABI.

Bar: bit-identical to nbhip_accelerate's tiled pass on a context holding the same bodies
(the same kernels and launch shape), on consecutive calls of different sizes -- the second
smaller, so stale partial sums or a stale radius flag in the shared scratch would show --
with and without radius cutoffs and softening; within 2e-6 of nbhip_accelerate's default,
which above 65,536 bodies is the pair-symmetric pass; and refusals that leave the GPU
untouched.
"""
import os

import ctypes as C

import numpy as np
import pytest

from nbody_amd import _lib, ics

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    # the HIP runtime the library itself links (one runtime per process)
    _lib.load()
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


class DevArray:
    """Caller-owned device memory, allocated before the entry point is called."""

    def __init__(self, L, nbytes):
        self.L = L
        self.ptr = C.c_void_p()
        assert L.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.L.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype=np.float32):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.L.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def split(b):
    pm = np.ascontiguousarray(b[:, [0, 1, 2, 7]], np.float32)
    vr = np.ascontiguousarray(b[:, [4, 5, 6, 3]], np.float32)
    return pm, vr


def reference(b, p, tiled=True):
    # nbhip_accelerate on a fresh (untuned) context holding the same bodies: its tiled
    # pass (the kernels the no-allocation entry point shares), or its default choice
    old = os.environ.get("NBHIP_DIRECT_SYM")
    if tiled:
        os.environ["NBHIP_DIRECT_SYM"] = "0"
    try:
        with _lib.Context(0) as c:
            c.upload_bodies(b)
            c.accelerate(p)
            return c.download(_lib.RB_ACCELERATIONS)[2]
    finally:
        if old is None:
            os.environ.pop("NBHIP_DIRECT_SYM", None)
        else:
            os.environ["NBHIP_DIRECT_SYM"] = old


CASES = [
    # (n, eps, radius) -- the second of each pair smaller than the first
    (30000, 0.01, 0.0), (5000, 0.01, 0.0),        # softened; j-split 2 then 8
    (20000, 0.0, 0.0), (3000, 0.02, 0.01),         # reference law; then cutoffs + softening
    (4096, 0.01, 0.02), (700, 0.01, 0.0),          # a radius flag set, then clear again
    (300000, 0.005, 0.0), (1, 0.01, 0.0),
]


def test_noalloc_matches_accelerate_on_consecutive_calls():
    L = hip()
    nmax = max(n for n, _, _ in CASES)
    pm_d, vr_d, acc_d = (DevArray(L, 16 * nmax) for _ in range(3))
    try:
        with _lib.Context(0) as c:
            for k, (n, eps, radius) in enumerate(CASES):
                print(f"noalloc case {k}: n={n} eps={eps} radius={radius}", flush=True)
                b = ics.plummer(n, seed=700 + k)
                b[:, 3] = radius
                p = _lib.params(eps=eps, G=[1.0, 2.0][k % 2])
                pm, vr = split(b)
                pm_d.put(pm)
                vr_d.put(vr)
                acc_d.put(np.full((n, 4), np.nan, np.float32))
                c.accelerate_noalloc(pm_d.ptr.value, vr_d.ptr.value, acc_d.ptr.value, n, p)
                c.synchronize()
                got = acc_d.get((n, 4))
                want = reference(b, p)
                assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), k
                if n >= 65536:
                    dflt = reference(b, p, tiled=False)[:, :3].astype(np.float64)
                    e = np.linalg.norm(got[:, :3] - dflt, axis=1) / np.linalg.norm(dflt, axis=1)
                    assert e.max() <= 2e-6, (k, e.max())
    finally:
        for d in (pm_d, vr_d, acc_d):
            d.free()


def test_noalloc_refusals():
    L = hip()
    n = 1000
    d = DevArray(L, 16 * n)
    try:
        with _lib.Context(0) as c:
            p = _lib.params(eps=0.01)
            host = np.zeros((n, 4), np.float32)
            bad = [
                (lambda: c.accelerate_noalloc(d.ptr.value, d.ptr.value, d.ptr.value, n,
                                              _lib.params(mode=_lib.MODE_NLOGN, size=64.0)), _lib.ERR_UNSUPPORTED),
                (lambda: c.accelerate_noalloc(d.ptr.value, d.ptr.value, d.ptr.value, _lib.MAX_BODIES + 1, p),
                 _lib.ERR_INVALID),
                (lambda: c.accelerate_noalloc(None, d.ptr.value, d.ptr.value, n, p), _lib.ERR_INVALID),
                (lambda: c.accelerate_noalloc(d.ptr.value + 4, d.ptr.value, d.ptr.value, n - 1, p), _lib.ERR_INVALID),
                (lambda: c.accelerate_noalloc(host.ctypes.data, d.ptr.value, d.ptr.value, n, p), _lib.ERR_INVALID),
                (lambda: c.accelerate_noalloc(d.ptr.value, d.ptr.value, d.ptr.value, n + 1, p), _lib.ERR_INVALID),
            ]
            for call, status in bad:
                with pytest.raises(_lib.HipError) as e:
                    call()
                assert e.value.status == status
            c.accelerate_noalloc(d.ptr.value, d.ptr.value, d.ptr.value, 0, p)   # nothing to do
            c.synchronize()
    finally:
        d.free()
