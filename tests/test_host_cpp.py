"""Runs the C++ conformance tests of the Solver layer (tests/cpp/test_hip_solver.cpp).

The binary links the C++ host layer (nbody_amd/libnbody_host.so) over the C ABI
(nbody_amd/libnbody_hip.so); each case restates a reference Catch2 test against the HIP
variants.  One process per tag.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_hip_solver")


def run(tag, env=None, timeout=300):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (python -m nbody_amd.build)")
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([BIN, tag], capture_output=True, text=True, timeout=timeout, env=e)
    print(r.stdout)
    return r


def test_unavailable_device_is_reported_not_crashed():
    # forced from the host, so it runs the same with or without a GPU
    r = run("cpu", {"NBODY_HIP_FORCE_UNAVAILABLE": "test"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[ PASS ] an unavailable variant is a no-op with a reason" in r.stdout
    assert "[ FAIL ]" not in r.stdout


@pytest.mark.gpu
def test_solver_conformance_on_device():
    r = run("gpu")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "[ FAIL ]" not in r.stdout
    assert " passed, 0 failed" in r.stdout
