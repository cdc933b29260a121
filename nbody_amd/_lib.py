"""ctypes binding of libnbody_hip.so (the C ABI declared in include/nbody_hip.h).

This module is plumbing for Python callers (tests, bench.py, the serving loop): it loads
the in-tree library and fails loudly when it is missing.  There is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libnbody_hip.so")

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_NOMEM = 3
ERR_UNSUPPORTED = 4

MODE_N2 = 0
MODE_NLOGN = 1
OPEN_REFERENCE = 0
OPEN_TEXTBOOK = 1

MAX_BODIES = 2147418112         # NBHIP_MAX_BODIES, 2^31 - 2^16
MAX_NLOGN_BODIES = (1 << 30) - 1  # NBHIP_MAX_NLOGN_BODIES

RB_NONE = 0
RB_POSITIONS = 1
RB_VELOCITIES = 2
RB_ACCELERATIONS = 4
RB_ALL = 7


BH_NODE_DTYPE = np.dtype([
    ("com", np.float32, 3), ("mass", np.float32), ("next", np.uint32),
    ("open_sq", np.float32), ("first", np.uint32), ("count", np.uint32),
])
assert BH_NODE_DTYPE.itemsize == 32

# The reference's 48-byte bh::Node (nbhip_tree_node).
TREE_NODE_DTYPE = np.dtype([
    ("center", np.float32, 3), ("size", np.float32),
    ("com", np.float32, 3), ("mass", np.float32),
    ("next", np.uint32), ("children", np.uint32), ("pad0", np.uint32), ("pad1", np.uint32),
])
assert TREE_NODE_DTYPE.itemsize == 48


class Params(C.Structure):
    """nbhip_params (include/nbody_hip.h)."""
    _fields_ = [
        ("dt", C.c_float), ("theta", C.c_float), ("G", C.c_float), ("size", C.c_float),
        ("wrap", C.c_int), ("mode", C.c_int), ("eps", C.c_float), ("opening", C.c_int),
        ("leaf_size", C.c_int),
    ]


BATCH_LAUNCH_GRID = 0
BATCH_LAUNCH_PERSISTENT = 1
BATCH_LAUNCH_PAIRS = 2

BATCH_GRAPH = 1
BATCH_PER_SYSTEM = 2
BATCH_ACCELERATE = 4
BATCH_COUNT = 8

BH_BATCH_MAX_SYSTEMS = 512
BH_BATCH_LEVELS = 18


class TuneEntry(C.Structure):
    """nbhip_tune_entry (include/nbody_hip.h)."""
    _fields_ = [("P", C.c_int), ("W", C.c_int), ("jsplit", C.c_int), ("ms", C.c_float)]


class BatchInfo(C.Structure):
    """nbhip_batch_info (include/nbody_hip.h)."""
    _fields_ = [
        ("systems", C.c_size_t), ("bodies", C.c_size_t), ("items", C.c_size_t),
        ("interactions_per_step", C.c_double), ("device_bytes", C.c_size_t),
        ("P", C.c_int), ("W", C.c_int), ("launch", C.c_int),
    ]


class BhBatchInfo(C.Structure):
    """nbhip_bh_batch_info (include/nbody_hip.h)."""
    _fields_ = [
        ("systems", C.c_size_t), ("bodies", C.c_size_t), ("device_bytes", C.c_size_t),
        ("levels", C.c_int), ("max_systems", C.c_int),
    ]


IC_PLUMMER = 0
IC_DISK_BULGE = 1
IC_DISK = 2
IC_CUBE = 3


class Ics(C.Structure):
    """nbhip_ics (include/nbody_hip.h): an initial-condition generator spec."""
    _fields_ = [("kind", C.c_int), ("reserved", C.c_int), ("seed", C.c_uint64),
                ("center", C.c_float * 3), ("vel", C.c_float * 3), ("axis", C.c_float * 3),
                ("G", C.c_float), ("mass", C.c_float), ("star_mass", C.c_float),
                ("scale", C.c_float), ("inner", C.c_float), ("thickness", C.c_float),
                ("z0", C.c_float), ("disk_frac", C.c_float), ("bulge_a", C.c_float),
                ("rcut", C.c_float)]


# Every symbol the header declares, with (restype, argtypes).
_vp = C.c_void_p
_sz = C.c_size_t
SIGNATURES = {
    "nbhip_default_params": (None, [C.POINTER(Params)]),
    "nbhip_probe": (C.c_int, [C.c_char_p, _sz]),
    "nbhip_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "nbhip_destroy": (None, [_vp]),
    "nbhip_last_error": (C.c_char_p, [_vp]),
    "nbhip_reserve": (C.c_int, [_vp, _sz]),
    "nbhip_count": (_sz, [_vp]),
    "nbhip_upload": (C.c_int, [_vp, _vp, _vp, _vp, _sz, _sz]),
    "nbhip_download": (C.c_int, [_vp, C.c_uint, _vp, _vp, _vp, _sz, _sz]),
    "nbhip_upload_bodies": (C.c_int, [_vp, _vp, _sz]),
    "nbhip_download_bodies": (C.c_int, [_vp, _vp, _sz]),
    "nbhip_accelerate": (C.c_int, [_vp, C.POINTER(Params)]),
    "nbhip_integrate": (C.c_int, [_vp, C.POINTER(Params)]),
    "nbhip_accelerate_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _sz, C.POINTER(Params)]),
    "nbhip_step": (C.c_int, [_vp, C.POINTER(Params), C.c_int]),
    "nbhip_step_graph": (C.c_int, [_vp, C.POINTER(Params), C.c_int]),
    "nbhip_synchronize": (C.c_int, [_vp]),
    "nbhip_test_raise_sort_fault": (C.c_int, [_vp]),
    "nbhip_query": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "nbhip_readback": (C.c_int, [_vp]),
    "nbhip_collect_bodies": (C.c_int, [_vp, _vp, _sz, C.POINTER(C.c_int)]),
    "nbhip_tune": (C.c_int, [_vp, C.POINTER(Params), _vp, _sz, C.POINTER(_sz), C.POINTER(TuneEntry)]),
    "nbhip_bh_tune": (C.c_int, [_vp, C.POINTER(Params), _vp, _sz, _vp, C.POINTER(C.c_int)]),
    "nbhip_bh_configure": (C.c_int, [_vp, C.c_int]),
    "nbhip_bh_root": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "nbhip_bh_sort_passes": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "nbhip_bh_profile": (C.c_int, [_vp, C.POINTER(Params), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "nbhip_diagnostics": (C.c_int, [_vp, C.POINTER(Params), _vp]),
    "nbhip_force_error_f64": (C.c_int, [_vp, C.POINTER(Params), _vp, _sz, _vp, _vp]),
    "nbhip_bh_keys": (C.c_int, [_vp, C.POINTER(Params), _vp, _vp]),
    "nbhip_bh_nodes": (C.c_int, [_vp, C.POINTER(Params), _vp, _sz, C.POINTER(_sz)]),
    "nbhip_bh_interactions": (C.c_int, [_vp, C.POINTER(Params), _vp, C.POINTER(C.c_uint64)]),
    "nbhip_bh_tree": (C.c_int, [_vp, C.POINTER(Params), _vp, _sz, C.POINTER(_sz)]),
    "nbhip_default_ics": (C.c_int, [C.c_int, C.POINTER(Ics)]),
    "nbhip_generate": (C.c_int, [_vp, C.POINTER(Ics), _sz]),
    "nbhip_batch_add_generated": (C.c_int, [_vp, C.POINTER(Ics), _sz, C.POINTER(Params), C.POINTER(C.c_int)]),
    "nbhip_batch_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "nbhip_batch_destroy": (None, [_vp]),
    "nbhip_batch_add": (C.c_int, [_vp, _vp, _sz, C.POINTER(Params), C.POINTER(C.c_int)]),
    "nbhip_batch_remove": (C.c_int, [_vp, C.c_int]),
    "nbhip_batch_step": (C.c_int, [_vp, C.c_int, C.c_uint]),
    "nbhip_batch_download": (C.c_int, [_vp, C.c_int, _vp, _sz]),
    "nbhip_batch_query": (C.c_int, [_vp, C.POINTER(BatchInfo)]),
    "nbhip_batch_configure": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "nbhip_batch_tune": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "nbhip_bh_batch_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "nbhip_bh_batch_destroy": (None, [_vp]),
    "nbhip_bh_batch_add": (C.c_int, [_vp, _vp, _sz, C.POINTER(Params), C.POINTER(C.c_int)]),
    "nbhip_bh_batch_add_generated": (C.c_int, [_vp, C.POINTER(Ics), _sz, C.POINTER(Params), C.POINTER(C.c_int)]),
    "nbhip_bh_batch_remove": (C.c_int, [_vp, C.c_int]),
    "nbhip_bh_batch_step": (C.c_int, [_vp, C.c_int, C.c_uint]),
    "nbhip_bh_batch_idle": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "nbhip_bh_batch_download": (C.c_int, [_vp, C.c_int, _vp, _sz]),
    "nbhip_bh_batch_visits": (C.c_int, [_vp, C.c_int, _vp, _sz]),
    "nbhip_bh_batch_query": (C.c_int, [_vp, C.POINTER(BhBatchInfo)]),
    "nbhip_bh_batch_profile": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "nbhip_timer_start": (C.c_int, [_vp]),
    "nbhip_timer_stop": (C.c_int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_long)]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (no GPU call is made).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m nbody_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class HipError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"nbhip status {status}: {msg}")
        self.status = status


def params(**kw) -> Params:
    p = Params()
    load().nbhip_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def ics(kind: int, **kw) -> Ics:
    """A generator spec: the kind's defaults (nbhip_default_ics), then `kw`."""
    s = Ics()
    if load().nbhip_default_ics(kind, C.byref(s)) != OK:
        raise ValueError(f"unknown initial-condition kind {kind}")
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        if k in ("center", "vel", "axis"):
            getattr(s, k)[:] = [float(x) for x in v]
        else:
            setattr(s, k, v)
    return s


def probe() -> tuple[bool, str]:
    buf = C.create_string_buffer(256)
    st = load().nbhip_probe(buf, 256)
    return st == OK, buf.value.decode()


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _f4(a, n) -> np.ndarray | None:
    if a is None:
        return None
    a = np.ascontiguousarray(a, np.float32)
    if a.size != 4 * n:
        raise ValueError(f"expected {4 * n} floats, got {a.size}")
    return a


class Context:
    """One device context: body buffers + a HIP stream (include/nbody_hip.h)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        st = self.L.nbhip_create(device, C.byref(h))
        if st != OK:
            raise HipError(st, self.L.nbhip_last_error(None).decode())
        self.h = h

    def _check(self, st: int) -> None:
        if st != OK:
            raise HipError(st, self.L.nbhip_last_error(self.h).decode())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nbhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- storage ----
    @property
    def count(self) -> int:
        return self.L.nbhip_count(self.h)

    def reserve(self, n: int) -> None:
        self._check(self.L.nbhip_reserve(self.h, n))

    def upload(self, pos_mass=None, vel_radius=None, acc=None, offset: int = 0, n: int | None = None):
        if n is None:
            for a in (pos_mass, vel_radius, acc):
                if a is not None:
                    n = np.asarray(a).size // 4
                    break
            else:
                n = 0
        pm, vr, ac = _f4(pos_mass, n), _f4(vel_radius, n), _f4(acc, n)
        self._check(self.L.nbhip_upload(self.h, _ptr(pm), _ptr(vr), _ptr(ac), offset, n))

    def download(self, mask: int = RB_ALL, offset: int = 0, n: int | None = None):
        if n is None:
            n = self.count - offset
        pm = np.zeros((n, 4), np.float32) if mask & RB_POSITIONS else None
        vr = np.zeros((n, 4), np.float32) if mask & RB_VELOCITIES else None
        ac = np.zeros((n, 4), np.float32) if mask & RB_ACCELERATIONS else None
        self._check(self.L.nbhip_download(self.h, mask, _ptr(pm), _ptr(vr), _ptr(ac), offset, n))
        return pm, vr, ac

    def upload_bodies(self, bodies: np.ndarray) -> None:
        b = np.ascontiguousarray(bodies, np.float32)
        if b.ndim != 2 or b.shape[1] != 12:
            raise ValueError("bodies must have shape (n, 12)")
        self._check(self.L.nbhip_upload_bodies(self.h, _ptr(b), b.shape[0]))

    def readback(self) -> None:
        """Start copying the bodies to pinned host memory behind the enqueued work."""
        self._check(self.L.nbhip_readback(self.h))

    def collect_bodies(self) -> tuple[np.ndarray, bool]:
        """The bodies (from the last readback when still current) and whether it was used."""
        b = np.zeros((self.count, 12), np.float32)
        st = C.c_int()
        self._check(self.L.nbhip_collect_bodies(self.h, _ptr(b), b.shape[0], C.byref(st)))
        return b, bool(st.value)

    def generate(self, spec: Ics, n: int) -> None:
        """Replace the bodies by n drawn on the device from `spec`."""
        self._check(self.L.nbhip_generate(self.h, C.byref(spec), n))

    def download_bodies(self, n: int | None = None) -> np.ndarray:
        if n is None:
            n = self.count
        b = np.zeros((n, 12), np.float32)
        self._check(self.L.nbhip_download_bodies(self.h, _ptr(b), n))
        return b

    # ---- stepping ----
    def accelerate(self, p: Params) -> None:
        self._check(self.L.nbhip_accelerate(self.h, C.byref(p)))

    def accelerate_noalloc(self, pos_mass_dev: int, vel_radius_dev: int, acc_dev: int, n: int, p: Params) -> None:
        """Direct-sum accelerations of n bodies in the caller's device arrays (addresses),
        enqueued on this context's stream without any allocation."""
        self._check(self.L.nbhip_accelerate_noalloc(self.h, pos_mass_dev, vel_radius_dev, acc_dev, n, C.byref(p)))

    def integrate(self, p: Params) -> None:
        self._check(self.L.nbhip_integrate(self.h, C.byref(p)))

    def step(self, p: Params, steps: int = 1) -> None:
        self._check(self.L.nbhip_step(self.h, C.byref(p), steps))

    def step_graph(self, p: Params, steps: int) -> None:
        """`steps` steps replayed from a captured hipGraph (captured on first use)."""
        self._check(self.L.nbhip_step_graph(self.h, C.byref(p), steps))

    def synchronize(self) -> None:
        self._check(self.L.nbhip_synchronize(self.h))

    def raise_sort_fault(self) -> None:
        """Testing hook: flag a radix-sort give-up from the host (no kernel runs)."""
        self._check(self.L.nbhip_test_raise_sort_fault(self.h))

    def idle(self) -> bool:
        flag = C.c_int()
        self._check(self.L.nbhip_query(self.h, C.byref(flag)))
        return bool(flag.value)

    # ---- autotuner ----
    def tune(self, p: Params) -> tuple[list[tuple[int, int, int, float]], tuple[int, int, int, float]]:
        """Time the launch shapes at the current N; returns (all timings, chosen)."""
        cap = 256
        arr = (TuneEntry * cap)()
        cnt = C.c_size_t()
        best = TuneEntry()
        self._check(self.L.nbhip_tune(self.h, C.byref(p), arr, cap, C.byref(cnt), C.byref(best)))
        rows = [(e.P, e.W, e.jsplit, e.ms) for e in arr[:cnt.value]]
        return rows, (best.P, best.W, best.jsplit, best.ms)

    BH_TUNE_DEFAULT = (1, 2, 4, 8, 12, 16, 24, 32)

    def bh_tune(self, p: Params, leaf_sizes=None) -> tuple[list[int], np.ndarray, int]:
        """Time a Barnes-Hut force pass at each leaf size (None: the library's defaults);
        returns (leaf sizes, ms per pass, fastest leaf size)."""
        best = C.c_int()
        if leaf_sizes is None:
            ms = np.zeros(len(self.BH_TUNE_DEFAULT), np.float32)
            self._check(self.L.nbhip_bh_tune(self.h, C.byref(p), None, ms.size, _ptr(ms), C.byref(best)))
            return list(self.BH_TUNE_DEFAULT), ms, best.value
        ls = np.ascontiguousarray(leaf_sizes, np.int32)
        ms = np.zeros(ls.size, np.float32)
        self._check(self.L.nbhip_bh_tune(self.h, C.byref(p), _ptr(ls), ls.size, _ptr(ms), C.byref(best)))
        return [int(v) for v in ls], ms, best.value

    def bh_profile(self, p: Params, steps: int) -> tuple[float, float]:
        """`steps` fused Barnes-Hut steps; mean (build ms, walk ms) from HIP events."""
        b, w = C.c_double(), C.c_double()
        self._check(self.L.nbhip_bh_profile(self.h, C.byref(p), steps, C.byref(b), C.byref(w)))
        return b.value, w.value

    def bh_sort_passes(self) -> int:
        """Radix passes the last Barnes-Hut key sort ran (8: a full sort; 0: none yet)."""
        v = C.c_int()
        self._check(self.L.nbhip_bh_sort_passes(self.h, C.byref(v)))
        return v.value

    def bh_root(self) -> float:
        """The root-cube edge of the last Barnes-Hut build (found from the bodies when the
        params' size was 0)."""
        v = C.c_float()
        self._check(self.L.nbhip_bh_root(self.h, C.byref(v)))
        return v.value

    def bh_configure(self, walk_split: int = 0) -> None:
        """Waves per 64-body group in the Barnes-Hut walk: 1, 2, 4, 8, or 0 for by body count."""
        self._check(self.L.nbhip_bh_configure(self.h, int(walk_split)))

    # ---- diagnostics ----
    def diagnostics(self, p: Params) -> np.ndarray:
        out = np.zeros(9, np.float64)
        self._check(self.L.nbhip_diagnostics(self.h, C.byref(p), _ptr(out)))
        return out

    def force_error_f64(self, p: Params, sample) -> tuple[np.ndarray, np.ndarray]:
        s = np.ascontiguousarray(sample, np.uint32)
        acc64 = np.zeros((s.size, 3), np.float64)
        rel = np.zeros(s.size, np.float64)
        self._check(self.L.nbhip_force_error_f64(self.h, C.byref(p), _ptr(s), s.size,
                                                 _ptr(acc64), _ptr(rel)))
        return acc64, rel

    # ---- Barnes-Hut internals ----
    def bh_keys(self, p: Params) -> tuple[np.ndarray, np.ndarray]:
        n = self.count
        keys = np.zeros(n, np.uint64)
        perm = np.zeros(n, np.uint32)
        self._check(self.L.nbhip_bh_keys(self.h, C.byref(p), _ptr(keys), _ptr(perm)))
        return keys, perm

    def bh_nodes(self, p: Params) -> np.ndarray:
        cnt = C.c_size_t()
        self._check(self.L.nbhip_bh_nodes(self.h, C.byref(p), None, 0, C.byref(cnt)))
        nodes = np.zeros(cnt.value, BH_NODE_DTYPE)
        self._check(self.L.nbhip_bh_nodes(self.h, C.byref(p), _ptr(nodes), nodes.size, C.byref(cnt)))
        return nodes

    def bh_tree(self, p: Params | None = None) -> np.ndarray:
        """The octree in the reference's bh::Node layout: built from the current positions
        with p, or (p None) the tree of the last Barnes-Hut build."""
        pp = C.byref(p) if p is not None else None
        cnt = C.c_size_t()
        self._check(self.L.nbhip_bh_tree(self.h, pp, None, 0, C.byref(cnt)))
        nodes = np.zeros(cnt.value, TREE_NODE_DTYPE)
        self._check(self.L.nbhip_bh_tree(self.h, pp, _ptr(nodes), nodes.size, C.byref(cnt)))
        return nodes

    def bh_interactions(self, p: Params) -> tuple[np.ndarray, int]:
        per = np.zeros(self.count, np.uint32)
        tot = C.c_uint64()
        self._check(self.L.nbhip_bh_interactions(self.h, C.byref(p), _ptr(per), C.byref(tot)))
        return per, tot.value

    # ---- timing ----
    def timer_start(self) -> None:
        self._check(self.L.nbhip_timer_start(self.h))

    def timer_stop(self) -> tuple[float, int]:
        ms = C.c_float()
        launches = C.c_long()
        self._check(self.L.nbhip_timer_stop(self.h, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value


class Batch:
    """Many independent direct-sum systems stepped together on one context."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.L = ctx.L
        h = C.c_void_p()
        ctx._check(self.L.nbhip_batch_create(ctx.h, C.byref(h)))
        self.h = h
        self.sizes: dict[int, int] = {}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nbhip_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, bodies: np.ndarray, p: Params) -> int:
        b = np.ascontiguousarray(bodies, np.float32)
        if b.ndim != 2 or b.shape[1] != 12:
            raise ValueError("bodies must have shape (n, 12)")
        i = C.c_int()
        self.ctx._check(self.L.nbhip_batch_add(self.h, _ptr(b), b.shape[0], C.byref(p), C.byref(i)))
        self.sizes[i.value] = b.shape[0]
        return i.value

    def add_generated(self, spec: Ics, n: int, p: Params) -> int:
        i = C.c_int()
        self.ctx._check(self.L.nbhip_batch_add_generated(self.h, C.byref(spec), n, C.byref(p), C.byref(i)))
        self.sizes[i.value] = n
        return i.value

    def remove(self, sid: int) -> None:
        self.ctx._check(self.L.nbhip_batch_remove(self.h, sid))
        self.sizes.pop(sid, None)

    def step(self, steps: int = 1, flags: int = 0) -> None:
        self.ctx._check(self.L.nbhip_batch_step(self.h, steps, flags))

    def download(self, sid: int) -> np.ndarray:
        b = np.zeros((self.sizes[sid], 12), np.float32)
        self.ctx._check(self.L.nbhip_batch_download(self.h, sid, _ptr(b), b.shape[0]))
        return b

    def info(self) -> BatchInfo:
        out = BatchInfo()
        self.ctx._check(self.L.nbhip_batch_query(self.h, C.byref(out)))
        return out

    def configure(self, P: int, W: int, launch: int = BATCH_LAUNCH_PERSISTENT) -> None:
        """P bodies per lane, W waves per block, launch mode BATCH_LAUNCH_GRID, _PERSISTENT
        (True and False mean those two) or _PAIRS."""
        self.ctx._check(self.L.nbhip_batch_configure(self.h, P, W, int(launch)))

    def tune(self) -> float:
        ms = C.c_double()
        self.ctx._check(self.L.nbhip_batch_tune(self.h, C.byref(ms)))
        return ms.value


class BhBatch:
    """Many independent Barnes-Hut systems: one build and one walk per step for all of
    them, on the batch's own stream (nbhip_bh_batch; errors report through `ctx`)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.L = ctx.L
        h = C.c_void_p()
        ctx._check(self.L.nbhip_bh_batch_create(ctx.h, C.byref(h)))
        self.h = h
        self.sizes: dict[int, int] = {}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.nbhip_bh_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, bodies: np.ndarray, p: Params) -> int:
        b = np.ascontiguousarray(bodies, np.float32)
        if b.ndim != 2 or b.shape[1] != 12:
            raise ValueError("bodies must have shape (n, 12)")
        i = C.c_int()
        self.ctx._check(self.L.nbhip_bh_batch_add(self.h, _ptr(b), b.shape[0], C.byref(p), C.byref(i)))
        self.sizes[i.value] = b.shape[0]
        return i.value

    def add_generated(self, spec: Ics, n: int, p: Params) -> int:
        i = C.c_int()
        self.ctx._check(self.L.nbhip_bh_batch_add_generated(self.h, C.byref(spec), n, C.byref(p), C.byref(i)))
        self.sizes[i.value] = n
        return i.value

    def remove(self, sid: int) -> None:
        self.ctx._check(self.L.nbhip_bh_batch_remove(self.h, sid))
        self.sizes.pop(sid, None)

    def step(self, steps: int = 1, flags: int = 0) -> None:
        self.ctx._check(self.L.nbhip_bh_batch_step(self.h, steps, flags))

    def idle(self) -> bool:
        out = C.c_int()
        self.ctx._check(self.L.nbhip_bh_batch_idle(self.h, C.byref(out)))
        return bool(out.value)

    def download(self, sid: int) -> np.ndarray:
        b = np.zeros((self.sizes[sid], 12), np.float32)
        self.ctx._check(self.L.nbhip_bh_batch_download(self.h, sid, _ptr(b), b.shape[0]))
        return b

    def visits(self, sid: int) -> np.ndarray:
        """Per-body interaction counts of the last BATCH_COUNT step."""
        v = np.zeros(self.sizes[sid], np.uint32)
        self.ctx._check(self.L.nbhip_bh_batch_visits(self.h, sid, _ptr(v), v.shape[0]))
        return v

    def info(self) -> BhBatchInfo:
        out = BhBatchInfo()
        self.ctx._check(self.L.nbhip_bh_batch_query(self.h, C.byref(out)))
        return out

    def profile(self, steps: int) -> tuple[float, float]:
        """`steps` fused steps timed by HIP events: mean (build_ms, walk_ms) per step."""
        b, w = C.c_double(), C.c_double()
        self.ctx._check(self.L.nbhip_bh_batch_profile(self.h, steps, C.byref(b), C.byref(w)))
        return b.value, w.value
