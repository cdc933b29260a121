"""The continuous-batching server (nbody_amd/serve.py) on the device.

Bars: every job completes; each job's result is bit-identical to the same job run alone
(direct-sum jobs in a batch of one, Barnes-Hut jobs on a context of their own); the
device-memory budget is never exceeded and a small budget still drains the queue.
"""
import numpy as np
import pytest

from nbody_amd import _lib, serve

pytestmark = pytest.mark.gpu


def alone(job):
    with _lib.Context(0) as ctx:
        bodies = job.bodies
        if bodies is None:  # drawn on the device: the same draw, by itself
            ctx.generate(job.ics, job.n)
            bodies = ctx.download_bodies()
        if job.mode == _lib.MODE_N2:
            with _lib.Batch(ctx) as b:
                sid = b.add(bodies, job.params)
                b.step(job.steps)
                return b.download(sid)
        ctx.upload_bodies(bodies)
        ctx.step(job.params, job.steps)
        return ctx.download_bodies()


def small_jobs(count=14, seed=5, device_ics=False):
    jobs = serve.make_jobs(count, seed=seed, n_range=(1024, 60000), steps_range=(3, 40),
                           direct_below=16384, device_ics=device_ics)
    serve.poisson_arrivals(jobs, rate=200.0, seed=seed)
    return jobs


@pytest.mark.parametrize("device_ics", [False, True])
def test_every_job_same_alone_and_in_the_mix(device_ics):
    jobs = small_jobs(device_ics=device_ics)
    assert any(j.mode == _lib.MODE_N2 for j in jobs) and any(j.mode == _lib.MODE_NLOGN for j in jobs)
    with serve.Server(chunk=8, bh_contexts=3, keep_results=True) as srv:
        rep = srv.run(jobs)
    s = rep.summary()
    print(s)
    assert rep.jobs == len(jobs) and s["latency_p99_s"] is not None
    assert rep.job_steps == sum(j.steps for j in jobs)
    for j in jobs:
        ref = alone(j)
        assert np.array_equal(j.result.view(np.uint32), ref.view(np.uint32)), (j.id, j.n, j.mode)


def test_budget_is_respected_and_queue_drains():
    # every job is there at the start, so the unbudgeted server admits at once every job
    # before the third Barnes-Hut one (two Barnes-Hut contexts, and the batch is idle); a
    # budget between the largest job and those jobs' bytes then has to hold jobs back
    def jobs():
        js = small_jobs(count=10, seed=5)
        for j in js:
            j.arrival = 0.0
        return js
    js = jobs()
    first, bh = 0, 0
    for j in js:
        bh += j.mode == _lib.MODE_NLOGN
        if bh > 2:
            break
        first += j.device_bytes()
    big = max(j.device_bytes() for j in js)
    assert first > big
    budget = (first + big) / 2
    with serve.Server(chunk=8, bh_contexts=2) as srv:
        free = srv.run(jobs())
    with serve.Server(hbm_budget=budget, chunk=8, bh_contexts=2) as srv:
        rep = srv.run(js)
    assert rep.jobs == len(js) and free.jobs == len(js)
    assert free.peak_device_bytes >= first > budget >= rep.peak_device_bytes
    # jobs held at once never exceed what the budget holds of the smallest job
    assert rep.max_concurrent <= int(budget // min(j.device_bytes() for j in js))


def test_rejects_impossible_jobs():
    jobs = small_jobs(count=2, seed=1)
    with serve.Server(hbm_budget=1000, chunk=8, bh_contexts=1) as srv:
        with pytest.raises(ValueError):
            srv.run(jobs)
