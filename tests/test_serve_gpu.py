"""The continuous-batching server (nbody_amd/serve.py) on the device.

Bars: every job completes; each job's result is bit-identical to the same job run alone
(direct-sum jobs in a batch of one; Barnes-Hut jobs in a Barnes-Hut batch of one, or on a
context of their own when the server pools contexts); the device-memory budget is never
exceeded and a small budget still drains the queue.
"""
import numpy as np
import pytest

from nbody_amd import _lib, serve

pytestmark = pytest.mark.gpu


MODES = ["batch", "contexts"]


def alone(job, bh_mode):
    with _lib.Context(0) as ctx:
        bodies = job.bodies
        if bodies is None:  # drawn on the device: the same draw, by itself
            ctx.generate(job.ics, job.n)
            bodies = ctx.download_bodies()
        if job.mode == _lib.MODE_N2:
            with _lib.Batch(ctx) as b:
                b.configure(*serve.DIRECT_BATCH_SHAPE)   # the server's launch shape
                sid = b.add(bodies, job.params)
                b.step(job.steps)
                return b.download(sid)
        if bh_mode == "batch":
            with _lib.BhBatch(ctx) as b:
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


@pytest.mark.parametrize("bh_mode", MODES)
@pytest.mark.parametrize("device_ics", [False, True])
def test_every_job_same_alone_and_in_the_mix(device_ics, bh_mode):
    jobs = small_jobs(device_ics=device_ics)
    assert any(j.mode == _lib.MODE_N2 for j in jobs) and any(j.mode == _lib.MODE_NLOGN for j in jobs)
    with serve.Server(chunk=8, bh_contexts=3, keep_results=True, bh_mode=bh_mode) as srv:
        rep = srv.run(jobs)
    s = rep.summary()
    print(s)
    assert rep.jobs == len(jobs) and s["latency_p99_s"] is not None
    assert rep.job_steps == sum(j.steps for j in jobs)
    for j in jobs:
        ref = alone(j, bh_mode)
        assert np.array_equal(j.result.view(np.uint32), ref.view(np.uint32)), (j.id, j.n, j.mode)


@pytest.mark.parametrize("bh_mode", MODES)
def test_budget_is_respected_and_queue_drains(bh_mode):
    # every job is there at the start, so the unbudgeted server admits at once every job
    # (both batches are idle), or with two Barnes-Hut contexts every job before the third
    # Barnes-Hut one; a budget between the largest job and those jobs' bytes then has to
    # hold jobs back
    def jobs():
        js = small_jobs(count=10, seed=5)
        for j in js:
            j.arrival = 0.0
        return js
    js = jobs()
    first, bh = 0, 0
    for j in js:
        bh += j.mode == _lib.MODE_NLOGN
        if bh > 2 and bh_mode == "contexts":
            break
        first += j.device_bytes()
    big = max(j.device_bytes() for j in js)
    assert first > big
    budget = (first + big) / 2
    with serve.Server(chunk=8, bh_contexts=2, bh_mode=bh_mode) as srv:
        free = srv.run(jobs())
    with serve.Server(hbm_budget=budget, chunk=8, bh_contexts=2, bh_mode=bh_mode) as srv:
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


@pytest.mark.parametrize("bh_mode", MODES)
def test_served_mix_at_workload_scale_matches_oracles(orc, bh_mode):
    # A served mix that reaches the workload's extremes (a Barnes-Hut job of 230,000 bodies
    # over 600 steps, direct jobs up to 1,000 steps), each job ending with a force pass at
    # its final positions.  Direct jobs: every job's forces on 256 sampled bodies against
    # the FP64 direct sum (1e-5).  Barnes-Hut jobs: the served forces equal a standalone
    # force pass on the same final bodies bit for bit (a Barnes-Hut batch of one, or a
    # context), and that pass matches the bucket oracle at the job's leaf size and octree
    # depth (18 levels in a batch, 21 on a context): >= 99.99% identical traversals, 1e-5
    # at 99.9%, 1e-4.
    jobs = serve.make_jobs(20, seed=17, device_ics=True)
    big = next(j for j in jobs if j.mode == _lib.MODE_NLOGN)
    big.n_bodies, big.steps = 230000, 600
    longest = next(j for j in jobs if j.mode == _lib.MODE_N2)
    longest.steps = 1000
    assert max(j.n for j in jobs) >= 200000 and max(j.steps for j in jobs) >= 500
    serve.poisson_arrivals(jobs, rate=40.0, seed=17)
    with serve.Server(chunk=16, bh_contexts=4, keep_results=True, final_forces=True, bh_mode=bh_mode) as srv:
        rep = srv.run(jobs)
    assert rep.jobs == len(jobs)
    rng = np.random.default_rng(17)
    with _lib.Context(0) as c:
        for j in jobs:
            res = j.result
            eps = float(j.params.eps)
            if j.mode == _lib.MODE_N2:
                s = rng.choice(j.n, min(j.n, 256), replace=False).astype(np.uint32)
                a64 = orc.direct_f64(res, s, G=float(j.params.G), eps=eps)
                e = np.linalg.norm(res[s, 8:11] - a64, axis=1) / np.linalg.norm(a64, axis=1)
                assert e.max() <= 1e-5, (j.id, j.n, e.max())
                continue
            c.upload_bodies(res)
            visits, _ = c.bh_interactions(j.params)
            fresh = c.download_bodies()
            root = c.bh_root()   # the jobs' roots are found from their bodies
            levels = 21
            if bh_mode == "batch":
                with _lib.BhBatch(c) as b:
                    sid = b.add(res, j.params)
                    b.step(1, _lib.BATCH_ACCELERATE | _lib.BATCH_COUNT)
                    fresh, visits, levels = b.download(sid), b.visits(sid), _lib.BH_BATCH_LEVELS
            assert np.array_equal(fresh[:, 8:11].view(np.uint32), res[:, 8:11].view(np.uint32)), j.id
            x = res.copy()
            _, xv = orc.bh_accelerate_buckets(x, root, theta=float(j.params.theta),
                                              leaf_size=int(j.params.leaf_size), eps=eps, max_level=levels)
            eq = visits == xv
            e = np.linalg.norm(fresh[:, 8:11] - x[:, 8:11], axis=1) / np.linalg.norm(x[:, 8:11], axis=1)
            print(f"job {j.id}: n={j.n} steps={j.steps} leaf={j.params.leaf_size} same traversal "
                  f"{eq.mean():.5f} 99.9% {np.quantile(e[eq], 0.999):.2e} max {e[eq].max():.2e}")
            assert eq.mean() >= 0.9999
            assert np.quantile(e[eq], 0.999) <= 1e-5 and e[eq].max() <= 1e-4
