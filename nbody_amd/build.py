"""Build recipe for libnbody_hip.so (HIP kernels + C ABI) and the C++ host layer.

    python -m nbody_amd.build          # compile everything in-tree

Every HIP translation unit is compiled for gfx950 only, with packed-FP32 instruction
selection switched off (-Xclang -target-feature -Xclang -packed-fp32-ops); hipcc's host
pass warns that it does not know that feature, which is harmless.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libnbody_hip.so")
HOSTLIB = os.path.join(PKG, "libnbody_host.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", INCLUDE, "-I", CSRC,
]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall", "-I", INCLUDE, "-I", HOST,
             "-I", os.path.join(HOST, "api")]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    err = "\n".join(ln for ln in r.stderr.splitlines()
                    if "packed-fp32-ops" not in ln and ln.strip())
    if "warning" in err or "error" in err:
        sys.stderr.write(err + "\n")


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", s, "-o", o]) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> str | None:
    """The C++ Solver layer (nbody_amd/host) over the C ABI, when present."""
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if not srcs:
        return None
    tmp = HOSTLIB + ".tmp"
    _run([CXX, *CXX_FLAGS, "-shared", "-o", tmp, *srcs, "-L", PKG, "-lnbody_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOSTLIB)
    return HOSTLIB


TESTS_CPP = os.path.join(ROOT, "tests", "cpp")
TEST_BIN = os.path.join(TESTS_CPP, "_build", "test_hip_solver")


def build_cpp_tests() -> str | None:
    """The C++ conformance tests of the Solver layer (tests/cpp)."""
    src = os.path.join(TESTS_CPP, "test_hip_solver.cpp")
    if not os.path.exists(src) or not os.path.exists(HOSTLIB):
        return None
    os.makedirs(os.path.dirname(TEST_BIN), exist_ok=True)
    tmp = TEST_BIN + ".tmp"
    _run([CXX, *CXX_FLAGS, "-o", tmp, src, "-L", PKG, "-lnbody_host", "-lnbody_hip",
          "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/../../../nbody_amd"])
    os.replace(tmp, TEST_BIN)
    return TEST_BIN


def build_all() -> None:
    build_hip()
    build_host()
    build_cpp_tests()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
