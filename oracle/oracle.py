"""ctypes wrapper over the CPU oracle (oracle/liboracle.so) and, when built, the reference's
own CPU sources (oracle/_ref/libnbref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker or the timed CPU baseline.  The product never imports it.

Bodies are numpy float32 arrays of shape (n, 12): the reference's 48-byte `nbody::Body`
(include/nbody/body.h:6-19) -- pos[0:3], radius[3], vel[4:7], mass[7], acc[8:11], pad[11].
Function-by-function citations of the reference live in nbody_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# The same source at -O3 -march=x86-64-v3 (identical results; the CPU baseline's second
# column).  variant("o3") makes the functions below use it.
LIB_O3_PATH = os.path.join(HERE, "liboracle_o3.so")
REF_PATH = os.path.join(HERE, "_ref", "libnbref.so")

NODE_DTYPE = np.dtype([
    ("center", np.float32, 3), ("size", np.float32),
    ("com", np.float32, 3), ("mass", np.float32),
    ("next", np.uint32), ("children", np.uint32), ("pad0", np.uint32), ("pad1", np.uint32),
])
assert NODE_DTYPE.itemsize == 48

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_nodep = np.ctypeslib.ndpointer(NODE_DTYPE, flags="C_CONTIGUOUS")

_lib = None
_libs: dict = {}
_path = LIB_PATH
_ref = None

# Worker threads when a caller passes threads=0: the host's cores, at most 16 (the CPU
# share of one GPU on the benchmark hosts).
DEFAULT_THREADS = max(1, min(16, os.cpu_count() or 1))


def _t(threads: int) -> int:
    return threads if threads > 0 else DEFAULT_THREADS


def build(ref: bool = True) -> None:
    """Compile liboracle.so and liboracle_o3.so (and _ref/libnbref.so when the reference
    tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so", "liboracle_o3.so"], check=True)
    if ref and os.path.isdir("/root/reference/source"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def variant(name: str = "o2") -> None:
    """Select the build the functions below call: "o2" (the default) or "o3"."""
    global _lib, _path
    _path = {"o2": LIB_PATH, "o3": LIB_O3_PATH}[name]
    _lib = _libs.get(_path)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_path):
        raise RuntimeError(f"oracle library not built: {_path} (run make -C oracle)")
    L = C.CDLL(_path)
    sz = C.c_size_t
    L.orc_hardware_threads.restype = C.c_int
    L.orc_gravity_pair.argtypes = [_f32p, C.c_float, _f32p, C.c_float, C.c_float, C.c_float, _f32p]
    L.orc_wrap.restype = C.c_float
    L.orc_wrap.argtypes = [C.c_float, C.c_float]
    L.orc_integrate.argtypes = [_f32p, sz, C.c_float, C.c_float, C.c_int, C.c_int]
    L.orc_brute_force.argtypes = [_f32p, sz, C.c_float, C.c_float, C.c_int]
    L.orc_brute_force_rows.argtypes = [_f32p, sz, _u32p, sz, C.c_float, C.c_float, C.c_int]
    L.orc_direct_f64.argtypes = [_f32p, sz, C.c_void_p, sz, C.c_double, C.c_double, _f64p, C.c_int]
    L.orc_energy_f64.argtypes = [_f32p, sz, C.c_double, C.c_double, _f64p, C.c_int]
    L.orc_morton_keys.argtypes = [_f32p, sz, C.c_float, C.c_int, _u64p]
    L.orc_stable_sort.argtypes = [_u64p, sz, _u32p]
    L.orc_bh_build.restype = sz
    L.orc_bh_build.argtypes = [_f32p, sz, C.c_float, _nodep, sz]
    L.orc_bh_accelerate.argtypes = [_f32p, sz, _nodep, sz, C.c_float, C.c_float, C.c_float,
                                    C.c_void_p, C.c_int]
    L.orc_bh_accelerate_f64com.argtypes = [_f32p, sz, _nodep, sz, C.c_float, C.c_double,
                                           C.c_double, C.c_void_p, C.c_int]
    L.orc_bh_accelerate_buckets.restype = C.c_int
    L.orc_bh_accelerate_buckets.argtypes = [_f32p, sz, _nodep, sz, C.c_float, C.c_int, C.c_int, C.c_int,
                                            C.c_double, C.c_double, _u64p, C.c_int]
    L.orc_node_size.restype = sz
    L.orc_philox4x64.argtypes = [_u64p, _u64p, _u64p]
    L.orc_generate.argtypes = [C.c_void_p, sz, _f32p, C.c_int]
    L.orc_generate.restype = C.c_int
    assert L.orc_node_size() == 48
    _lib = _libs[_path] = L
    return L


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own CPU sources (None when oracle/_ref was not built)."""
    global _ref
    if _ref is not None:
        return _ref
    if not ref_available():
        return None
    R = C.CDLL(REF_PATH)
    sz = C.c_size_t
    R.ref_gravity_pair.argtypes = [_f32p, C.c_float, _f32p, C.c_float, C.c_float, _f32p]
    R.ref_wrap.restype = C.c_float
    R.ref_wrap.argtypes = [C.c_float, C.c_float]
    R.ref_brute_force.argtypes = [_f32p, sz, C.c_float]
    R.ref_integrate.argtypes = [_f32p, sz, C.c_float, C.c_float, C.c_int]
    R.ref_build_tree.restype = sz
    R.ref_build_tree.argtypes = [_f32p, sz, C.c_float, _nodep, sz]
    R.ref_bh_accelerate.argtypes = [_f32p, sz, C.c_float, C.c_float, C.c_float, C.c_void_p]
    R.ref_cell_keys.argtypes = [_f32p, sz, C.c_float, C.c_int, _u64p]
    R.ref_disk.argtypes = [_f32p, sz, C.c_float]
    R.ref_cube.argtypes = [_f32p, sz]
    _ref = R
    return R


def _bodies(b: np.ndarray) -> np.ndarray:
    if b.dtype != np.float32 or b.ndim != 2 or b.shape[1] != 12 or not b.flags.c_contiguous:
        raise ValueError("bodies must be a C-contiguous float32 array of shape (n, 12)")
    return b


# ---- oracle ---------------------------------------------------------------------------

def hardware_threads() -> int:
    return lib().orc_hardware_threads()


def gravity_pair(pos, radius, src, src_mass, G=1.0, eps=0.0) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_gravity_pair(np.asarray(pos, np.float32), radius, np.asarray(src, np.float32),
                           src_mass, G, eps, out)
    return out


def wrap(x: float, size: float) -> float:
    return lib().orc_wrap(x, size)


def brute_force(bodies: np.ndarray, G=1.0, eps=0.0, threads=0) -> np.ndarray:
    """Accelerations in place (bodies[:, 8:11]); returns bodies."""
    b = _bodies(bodies)
    lib().orc_brute_force(b, b.shape[0], G, eps, _t(threads))
    return b


def brute_force_rows(bodies: np.ndarray, rows: np.ndarray, G=1.0, eps=0.0, threads=0) -> np.ndarray:
    b = _bodies(bodies)
    r = np.ascontiguousarray(rows, np.uint32)
    lib().orc_brute_force_rows(b, b.shape[0], r, r.size, G, eps, _t(threads))
    return b


def integrate(bodies: np.ndarray, dt: float, size=10000.0, wrap=True, threads=0) -> np.ndarray:
    b = _bodies(bodies)
    lib().orc_integrate(b, b.shape[0], dt, size, int(bool(wrap)), _t(threads))
    return b


def direct_f64(bodies: np.ndarray, sample=None, G=1.0, eps=0.0, threads=0) -> np.ndarray:
    """FP64 accelerations, shape (len(sample) or n, 3)."""
    b = _bodies(bodies)
    if sample is None:
        out = np.zeros((b.shape[0], 3), np.float64)
        lib().orc_direct_f64(b, b.shape[0], None, 0, G, eps, out, _t(threads))
    else:
        s = np.ascontiguousarray(sample, np.uint32)
        out = np.zeros((s.size, 3), np.float64)
        lib().orc_direct_f64(b, b.shape[0], s.ctypes.data, s.size, G, eps, out, _t(threads))
    return out


def energy_f64(bodies: np.ndarray, G=1.0, eps=0.0, threads=0) -> np.ndarray:
    """[kinetic, potential, px, py, pz, Lx, Ly, Lz, mass] in FP64."""
    b = _bodies(bodies)
    out = np.zeros(9, np.float64)
    lib().orc_energy_f64(b, b.shape[0], G, eps, out, _t(threads))
    return out


def morton_keys(bodies: np.ndarray, size: float, levels: int = 21) -> np.ndarray:
    b = _bodies(bodies)
    keys = np.zeros(b.shape[0], np.uint64)
    lib().orc_morton_keys(b, b.shape[0], size, levels, keys)
    return keys


def stable_sort(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64)
    perm = np.zeros(k.size, np.uint32)
    lib().orc_stable_sort(k, k.size, perm)
    return perm


def bh_build(bodies: np.ndarray, size: float) -> np.ndarray:
    b = _bodies(bodies)
    cap = max(64, 16 * b.shape[0] + 64)
    while True:
        nodes = np.zeros(cap, NODE_DTYPE)
        n = lib().orc_bh_build(b, b.shape[0], size, nodes, cap)
        if n:
            return nodes[:n].copy()
        cap *= 4


def bh_accelerate(bodies: np.ndarray, size: float, theta=0.5, G=1.0, eps=0.0, threads=0,
                  nodes=None, visits=False):
    """Reference Barnes-Hut accelerations in place.  Returns (bodies, visits or None)."""
    b = _bodies(bodies)
    if nodes is None:
        nodes = bh_build(b, size)
    v = np.zeros(b.shape[0], np.uint64) if visits else None
    lib().orc_bh_accelerate(b, b.shape[0], nodes, nodes.size, theta, G, eps,
                            v.ctypes.data if v is not None else None, _t(threads))
    return b, v


def bh_accelerate_f64com(bodies: np.ndarray, size: float, theta=0.5, G=1.0, eps=0.0, threads=0,
                         nodes=None, visits=False):
    """The reference tree and traversal with FP64 centres of mass and FP64 force sums
    (the acceptance test still in FP32).  Returns (bodies, visits or None)."""
    b = _bodies(bodies)
    if nodes is None:
        nodes = bh_build(b, size)
    v = np.zeros(b.shape[0], np.uint64) if visits else None
    lib().orc_bh_accelerate_f64com(b, b.shape[0], nodes, nodes.size, theta, G, eps,
                                   v.ctypes.data if v is not None else None, _t(threads))
    return b, v


def bh_accelerate_buckets(bodies: np.ndarray, size: float, theta=0.5, leaf_size=1, G=1.0, eps=0.0,
                          textbook=False, threads=0, nodes=None, max_level=21):
    """The reference tree and traversal with leaf buckets of at most `leaf_size` bodies
    (and cells max_level levels deep: 21 for one system, 18 in a batch of systems): a
    bucket is tested whole and, when opened, applied body by body.  FP64 centres and sums;
    the acceptance test in FP32.  Accelerations in place; returns (bodies, per-body visit
    counts)."""
    b = _bodies(bodies)
    if nodes is None:
        nodes = bh_build(b, size)
    v = np.zeros(b.shape[0], np.uint64)
    if lib().orc_bh_accelerate_buckets(b, b.shape[0], nodes, nodes.size, theta, int(bool(textbook)),
                                       int(leaf_size), int(max_level), G, eps, v, _t(threads)) != 0:
        raise MemoryError("orc_bh_accelerate_buckets")
    return b, v


# ---- the reference's own sources (oracle/_ref) ------------------------------------------

def ref_brute_force(bodies: np.ndarray, G=1.0) -> np.ndarray:
    b = _bodies(bodies)
    ref().ref_brute_force(b, b.shape[0], G)
    return b


def ref_integrate(bodies: np.ndarray, dt: float, size=10000.0, wrap=True) -> np.ndarray:
    b = _bodies(bodies)
    ref().ref_integrate(b, b.shape[0], dt, size, int(bool(wrap)))
    return b


def ref_build_tree(bodies: np.ndarray, size: float) -> np.ndarray:
    b = _bodies(bodies)
    cap = max(64, 16 * b.shape[0] + 64)
    while True:
        nodes = np.zeros(cap, NODE_DTYPE)
        n = ref().ref_build_tree(b, b.shape[0], size, nodes, cap)
        if n <= cap:
            return nodes[:n].copy()
        cap = n


def ref_bh_accelerate(bodies: np.ndarray, size: float, theta=0.5, G=1.0, visits=False):
    b = _bodies(bodies)
    v = np.zeros(b.shape[0], np.uint64) if visits else None
    ref().ref_bh_accelerate(b, b.shape[0], size, theta, G, v.ctypes.data if v is not None else None)
    return b, v


def ref_gravity_pair(pos, radius, src, src_mass, G=1.0) -> np.ndarray:
    out = np.zeros(3, np.float32)
    ref().ref_gravity_pair(np.asarray(pos, np.float32), radius, np.asarray(src, np.float32),
                           src_mass, G, out)
    return out


def ref_wrap(x: float, size: float) -> float:
    return ref().ref_wrap(x, size)


def ref_cell_keys(bodies: np.ndarray, size: float, levels: int = 21) -> np.ndarray:
    b = _bodies(bodies)
    keys = np.zeros(b.shape[0], np.uint64)
    ref().ref_cell_keys(b, b.shape[0], size, levels, keys)
    return keys


def ref_disk(n: int, outer_radius=100.0) -> np.ndarray:
    b = np.zeros((n, 12), np.float32)
    ref().ref_disk(b, n, outer_radius)
    return b


def ref_cube(n: int) -> np.ndarray:
    b = np.zeros((n, 12), np.float32)
    ref().ref_cube(b, n)
    return b


# ---- initial conditions ------------------------------------------------------------------

class Ics(C.Structure):
    """The generator spec, field for field the layout of the library's nbhip_ics."""
    _fields_ = [("kind", C.c_int), ("reserved", C.c_int), ("seed", C.c_uint64),
                ("center", C.c_float * 3), ("vel", C.c_float * 3), ("axis", C.c_float * 3),
                ("G", C.c_float), ("mass", C.c_float), ("star_mass", C.c_float),
                ("scale", C.c_float), ("inner", C.c_float), ("thickness", C.c_float),
                ("z0", C.c_float), ("disk_frac", C.c_float), ("bulge_a", C.c_float),
                ("rcut", C.c_float)]


def philox4x64(counter, key) -> np.ndarray:
    c = np.ascontiguousarray(counter, np.uint64)
    k = np.ascontiguousarray(key, np.uint64)
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64(c, k, out)
    return out


def generate(spec, n: int, threads=0) -> np.ndarray:
    """n reference Body records drawn from `spec` (any ctypes struct with Ics's layout)."""
    s = Ics.from_buffer_copy(bytes(spec))
    out = np.zeros((n, 12), np.float32)
    if lib().orc_generate(C.byref(s), n, out, _t(threads)) != 0:
        raise MemoryError("orc_generate")
    return out

