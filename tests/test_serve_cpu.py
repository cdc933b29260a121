"""Host-side pieces of the server that need no device: the workload generator and the
arrival process (nbody_amd/serve.py)."""
import numpy as np

from nbody_amd import _lib, serve


def test_workload_mix_and_modes():
    jobs = serve.make_jobs(40, seed=3, n_range=(1024, 4096 * 8), steps_range=(10, 1000),
                           direct_below=16384)
    ns = np.array([j.n for j in jobs])
    assert ns.min() >= 1024 and ns.max() < 4096 * 8
    assert all((j.mode == _lib.MODE_N2) == (j.n < 16384) for j in jobs)
    assert all(10 <= j.steps <= 1000 for j in jobs)
    assert {j.tag for j in jobs} == {"disk", "plummer"}
    for j in jobs:
        assert j.bodies.dtype == np.float32 and j.bodies.shape == (j.n, 12)
        if j.mode == _lib.MODE_NLOGN:
            assert np.abs(j.bodies[:, 0:3]).max() < j.params.size / 2   # inside the root cube


def test_poisson_arrivals():
    jobs = serve.make_jobs(2000, seed=1, n_range=(1024, 1100), steps_range=(10, 11))
    serve.poisson_arrivals(jobs, rate=50.0, seed=2)
    t = np.array([j.arrival for j in jobs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    gaps = np.diff(t)
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0     # mean inter-arrival 1/rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1     # exponential: cv = 1


def test_device_ics_jobs_carry_specs(orc):
    jobs = serve.make_jobs(12, seed=4, n_range=(1024, 60000), device_ics=True)
    host = serve.make_jobs(12, seed=4, n_range=(1024, 60000))
    for j, h in zip(jobs, host):
        assert j.bodies is None and j.n == h.n and j.mode == h.mode and j.steps == h.steps
        assert j.ics.kind == (_lib.IC_DISK_BULGE if j.tag == "disk" else _lib.IC_PLUMMER)
        assert j.device_bytes() == h.device_bytes()
    for j in jobs[:4]:
        b = orc.generate(j.ics, j.n)   # what the device draws, within the root cube
        assert np.abs(b[:, 0:3]).max() < j.params.size / 2

