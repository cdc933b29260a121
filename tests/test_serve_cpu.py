"""Host-side pieces of the server that need no device: the workload generator and the
arrival process (nbody_amd/serve.py)."""
import numpy as np
import pytest

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
            assert j.params.size == 0.0   # the root cube is found from the bodies


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
        b = orc.generate(j.ics, j.n)   # what the device draws, within the fixed cube of the direct jobs
        assert np.abs(b[:, 0:3]).max() < 64.0



class _FakeCtx:
    """Stands in for a device context: a chunk of k steps stays in flight for k polls."""

    def __init__(self, device=0):
        self.busy = 0
        self.n = 0

    def idle(self):
        if self.busy > 0:
            self.busy -= 1
            return False
        return True

    def generate(self, spec, n):
        self.n = n

    def upload_bodies(self, b):
        self.n = b.shape[0]

    def step(self, p, k):
        self.busy = k

    step_graph = step

    def download_bodies(self):
        return np.zeros((self.n, 12), np.float32)

    def bh_tune(self, p):
        return [16], np.array([0.1]), 16

    def close(self):
        pass


class _FakeBatch:
    def __init__(self, ctx):
        self.ctx, self.sizes, self.next = ctx, {}, 1

    def add_generated(self, spec, n, p):
        self.sizes[self.next] = n
        self.next += 1
        return self.next - 1

    def add(self, bodies, p):
        return self.add_generated(None, bodies.shape[0], p)

    def remove(self, sid):
        del self.sizes[sid]

    def step(self, k, flags=0):
        self.ctx.busy = k

    def download(self, sid):
        return np.zeros((self.sizes[sid], 12), np.float32)

    def configure(self, *a):
        pass

    def close(self):
        pass


class _FakeBhBatch(_FakeBatch):
    """The Barnes-Hut batch runs on its own stream: its chunk stays in flight for k polls."""

    def __init__(self, ctx):
        super().__init__(ctx)
        self.busy = 0

    def step(self, k, flags=0):
        self.busy = k

    def idle(self):
        if self.busy > 0:
            self.busy -= 1
            return False
        return True


def _fake_device(monkeypatch):
    monkeypatch.setattr(_lib, "Context", _FakeCtx)
    monkeypatch.setattr(_lib, "Batch", _FakeBatch)
    monkeypatch.setattr(_lib, "BhBatch", _FakeBhBatch)


def _admission_order(monkeypatch, policy, bh_mode):
    _fake_device(monkeypatch)
    jobs = serve.make_jobs(3, seed=1, n_range=(1024, 1025), steps_range=(64, 64), device_ics=True)
    d1, d2, b1 = jobs
    b1.params = _lib.params(mode=_lib.MODE_NLOGN, size=128.0, dt=1e-3, eps=0.01)
    b1.n_bodies = 20000
    d1.arrival, d2.arrival, b1.arrival = 0.0, 0.05, 0.06   # d2 and b1 arrive while d1's chunk runs
    ticks = iter(range(10 ** 7))
    with serve.Server(chunk=64, bh_contexts=1, admission=policy, bh_mode=bh_mode) as srv:
        rep = srv.run(jobs, clock=lambda: next(ticks) * 1e-3)
    assert rep.jobs == 3
    return d1, d2, b1


@pytest.mark.parametrize("bh_mode", ["batch", "contexts"])
def test_per_class_admission_does_not_block_across_classes(monkeypatch, bh_mode):
    # the batch is busy with d1 when d2 (direct) and b1 (Barnes-Hut) arrive: per class,
    # b1 joins the idle Barnes-Hut batch (or takes the free context) at once while d2 waits
    # for the chunk boundary; under one FCFS queue d2 at the head holds b1 back until then
    d1, d2, b1 = _admission_order(monkeypatch, "per_class", bh_mode)
    assert b1.admitted < d2.admitted
    d1, d2, b1 = _admission_order(monkeypatch, "fcfs", bh_mode)
    assert b1.admitted >= d2.admitted


@pytest.mark.parametrize("bh_mode", ["batch", "contexts"])
def test_budget_held_in_arrival_order(monkeypatch, bh_mode):
    # a direct job blocked by the busy batch keeps its place for the budget: a later
    # Barnes-Hut job that would only fit by taking that place waits
    _fake_device(monkeypatch)
    jobs = serve.make_jobs(3, seed=1, n_range=(1024, 1025), steps_range=(64, 64), device_ics=True)
    d1, d2, b1 = jobs
    b1.params = _lib.params(mode=_lib.MODE_NLOGN, size=128.0, dt=1e-3, eps=0.01)
    b1.n_bodies = 2048
    d1.arrival, d2.arrival, b1.arrival = 0.0, 0.05, 0.06
    budget = d1.device_bytes() + d2.device_bytes() + b1.device_bytes() - 1
    ticks = iter(range(10 ** 7))
    with serve.Server(hbm_budget=budget, chunk=64, bh_contexts=1, bh_mode=bh_mode) as srv:
        srv.run(jobs, clock=lambda: next(ticks) * 1e-3)
    assert b1.admitted > d2.admitted
