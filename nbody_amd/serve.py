"""Continuous-batching service for many simulation jobs on one MI355X.

Jobs arrive over time (each: bodies, a step count, a mode).  The server keeps every
admitted job's state on the device until its last step, and runs:
  * direct-sum jobs (N below `direct_below`) together in one nbhip_batch -- one launch per
    step for all of them, with jobs joining and leaving between chunks of steps;
  * Barnes-Hut jobs together in one nbhip_bh_batch (bh_mode "batch") -- one radix sort, one
    pass of each build kernel and one walk per step for all of them, on the batch's own
    stream, concurrent with the direct-sum batch, joining and leaving between chunks; or
    (bh_mode "contexts") each on its own context from a fixed pool.
Admission keeps one first-come-first-served queue per resource class (direct sum, Barnes-Hut):
the device-memory budget (bytes per body estimated per mode) is granted in arrival order
across both, but a class whose own resource is busy (the batch between chunk boundaries,
no free context) does not hold up the other class; a job that does not fit waits.
A job brings its bodies either as host records (uploaded at admission) or as a generator
spec (nbhip_ics), in which case its bodies are drawn on the device and never cross PCIe.
Each resource advances in chunks of at most `chunk` steps, never past the step count of
any job it holds, and is polled without blocking (nbhip_query), so a resource that
finishes early is refilled while the others still run.

A job's result does not depend on what else is running: a batched system sums each
body's force in a fixed order whatever shares the batch, and a Barnes-Hut system's sort
order, prefix sums, nodes and walk are its own in the Barnes-Hut batch (a pooled job owns
its context).  Latency is measured from a job's arrival to its result being back on the host.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque

import numpy as np

from . import _lib

# Device bytes per body, for admission: state arrays (pm x2, vr, acc: 64 B) for a
# batched job; for a Barnes-Hut job also keys, sorted copies, prefix sums and up to 2N
# nodes (about 170 B more) and the FP64 scratch of bodies drawn on the device (48 B) --
# in the Barnes-Hut batch the spare arrays a removal compacts into take the place of the
# second position buffer and the generator's scratch.
BYTES_PER_BODY = {_lib.MODE_N2: 64, _lib.MODE_NLOGN: 288}


@dataclasses.dataclass
class Job:
    bodies: np.ndarray | None       # reference Body records, float32 (n, 12), or None ...
    steps: int
    params: _lib.Params
    arrival: float = 0.0            # seconds after the start of the run
    tag: str = ""
    ics: _lib.Ics | None = None     # ... and the bodies are drawn on the device from this
    n_bodies: int = 0               # with this many bodies
    # filled in by the server
    id: int = -1
    admitted: float = float("nan")
    done: float = float("nan")
    result: np.ndarray | None = None
    interactions_per_step: float = 0.0

    @property
    def n(self) -> int:
        return self.bodies.shape[0] if self.bodies is not None else self.n_bodies

    @property
    def mode(self) -> int:
        return self.params.mode

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    def device_bytes(self) -> int:
        return BYTES_PER_BODY[self.mode] * self.n


@dataclasses.dataclass
class Report:
    jobs: int
    wall_s: float
    job_steps: int
    body_steps: float
    interactions: float
    latencies: np.ndarray
    peak_device_bytes: int
    max_concurrent: int

    def summary(self) -> dict:
        lat = self.latencies
        return {
            "jobs": self.jobs, "wall_s": self.wall_s,
            "jobs_per_s": self.jobs / self.wall_s,
            "steps_per_s": self.job_steps / self.wall_s,
            "body_steps_per_s": self.body_steps / self.wall_s,
            "interactions_per_s": self.interactions / self.wall_s,
            "latency_p50_s": float(np.percentile(lat, 50)) if lat.size else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if lat.size else None,
            "latency_max_s": float(lat.max()) if lat.size else None,
            "peak_device_bytes": self.peak_device_bytes,
            "max_concurrent_jobs": self.max_concurrent,
        }


class _BhSlot:
    def __init__(self):
        self.ctx = _lib.Context(0)
        self.job: Job | None = None
        self.remaining = 0
        self.inflight = 0


# The direct-sum batch's launch shape: pair-symmetric tiles of 512-body blocks (P = 2,
# W = 4).  On the served mix's direct-sum jobs (N in [1k, 16k)) a few at a time, as they
# run, it takes 0.15-0.28 ms per step for 4-16 systems against 0.29-0.42 for the batch's
# one-sided default (P = 2, W = 8, persistent), and 1.00 against 1.50 ms for all 75 at once.
DIRECT_BATCH_SHAPE = (2, 4, _lib.BATCH_LAUNCH_PAIRS)


class Server:
    def __init__(self, hbm_budget: float = 200e9, bh_contexts: int = 8, chunk: int = 16,
                 direct_below: int = 16384, keep_results: bool = False, batch_shape=DIRECT_BATCH_SHAPE,
                 bh_graph: bool = True, bh_leaf="auto", admission: str = "per_class",
                 count_bh_interactions: bool = False, final_forces: bool = False,
                 bh_mode: str = "batch"):
        self.budget = hbm_budget
        if bh_mode not in ("batch", "contexts"):
            raise ValueError(bh_mode)
        self.bh_mode = bh_mode
        # end each job with one accelerate-only pass, so its result carries the forces at
        # its final positions (a fused step leaves those of the positions before it); for
        # checking served results, it costs a force pass per job
        self.final_forces = final_forces
        # "per_class": a busy resource class does not block the other's queue; "fcfs": one
        # queue, the earliest waiting job blocks every later one (the earlier policy, kept
        # for comparison)
        if admission not in ("per_class", "fcfs"):
            raise ValueError(admission)
        self.admission = admission
        # count each Barnes-Hut job's body-node interactions with one extra force pass after
        # its result is back (off the admission path); off, Report.interactions covers the
        # direct-sum jobs only
        self.count_bh_interactions = count_bh_interactions
        # Barnes-Hut leaf size: a number for every job, or "auto": the autotuner
        # (nbhip_bh_tune) times the candidates on the first job of each power-of-two size
        # class, on that job's own bodies, and the class keeps its pick for the run.
        self.bh_leaf = bh_leaf
        self.leaf_by_class: dict[int, int] = {}
        self.chunk = chunk
        # Barnes-Hut chunks replayed from a captured graph (nbhip_step_graph, or the batch's
        # NBHIP_BATCH_GRAPH): one launch per chunk instead of ~20 per step; captured again
        # when a chunk's length or the batch's membership changes.
        self.bh_graph = bh_graph
        self.direct_below = direct_below
        self.keep_results = keep_results
        self.batch_ctx = _lib.Context(0)
        self.batch = _lib.Batch(self.batch_ctx)
        if batch_shape is not None:
            self.batch.configure(*batch_shape)
        self.slots = [_BhSlot() for _ in range(bh_contexts)] if bh_mode == "contexts" else []
        # the Barnes-Hut batch (its own stream), and a context for the leaf tuner and the
        # interaction counts, which run on one job's bodies
        self.bh_ctx = _lib.Context(0)
        self.bh = _lib.BhBatch(self.bh_ctx) if bh_mode == "batch" else None

    def close(self) -> None:
        self.batch.close()
        self.batch_ctx.close()
        if self.bh is not None:
            self.bh.close()
        self.bh_ctx.close()
        for s in self.slots:
            s.ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _leaf_for(self, job: Job, ctx: _lib.Context | None) -> int:
        """The job's bucket size; ctx holds its bodies (None: load them into the tuning
        context when the tuner needs them)."""
        if self.bh_leaf != "auto":
            return int(self.bh_leaf)
        cls = max(0, int(job.n).bit_length() - 1)
        if cls not in self.leaf_by_class:
            if ctx is None:
                ctx = self.bh_ctx
                self._load(ctx, job)
            # the accelerations the tuner leaves behind are recomputed by the job's first step
            _, _, best = ctx.bh_tune(job.params)
            self.leaf_by_class[cls] = best
        return self.leaf_by_class[cls]

    @staticmethod
    def _load(ctx: _lib.Context, job: Job) -> None:
        if job.bodies is not None:
            ctx.upload_bodies(job.bodies)
        else:
            ctx.generate(job.ics, job.n)

    def _count(self, job: Job, result: np.ndarray) -> float:
        """Body-node interactions of one force pass at the job's final positions."""
        self.bh_ctx.upload_bodies(result)
        _, inter = self.bh_ctx.bh_interactions(job.params)
        job.interactions_per_step = float(inter)
        return job.steps * job.interactions_per_step

    # ---- the loop ------------------------------------------------------------------------
    def run(self, jobs: list[Job], clock=time.perf_counter) -> Report:
        for k, j in enumerate(jobs):
            j.id = k
            if j.mode == _lib.MODE_N2 and j.n >= self.direct_below:
                raise ValueError(f"job {k}: direct sum is for N < {self.direct_below}")
            if j.device_bytes() > self.budget:
                raise ValueError(f"job {k} needs {j.device_bytes()} bytes, more than the budget")
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        waiting: deque[Job] = deque()
        direct: dict[int, tuple[Job, int]] = {}   # batch system id -> (job, steps remaining)
        batch_inflight = 0
        bh_jobs: dict[int, tuple[Job, int]] = {}  # Barnes-Hut batch system id -> (job, remaining)
        bh_inflight = 0
        in_use = 0
        peak = 0
        max_conc = 0
        job_steps = 0
        body_steps = 0.0
        interactions = 0.0
        finished: list[Job] = []
        t0 = clock()

        def now() -> float:
            return clock() - t0

        def finish(job: Job, result: np.ndarray) -> None:
            nonlocal in_use
            job.done = now()
            if self.keep_results:
                job.result = result
            in_use -= job.device_bytes()
            finished.append(job)

        while len(finished) < len(jobs):
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())

            # -- the batch: account a finished chunk, retire jobs, admit, launch --------
            if batch_inflight and self.batch_ctx.idle():
                if self.final_forces and any(rem <= batch_inflight for _, rem in direct.values()):
                    self.batch.step(1, _lib.BATCH_ACCELERATE)   # positions do not move
                for sid in list(direct):
                    job, rem = direct[sid]
                    rem -= batch_inflight
                    job_steps += batch_inflight
                    body_steps += batch_inflight * job.n
                    interactions += batch_inflight * float(job.n) ** 2
                    if rem <= 0:
                        res = self.batch.download(sid)
                        self.batch.remove(sid)
                        del direct[sid]
                        finish(job, res)
                    else:
                        direct[sid] = (job, rem)
                batch_inflight = 0

            # -- the Barnes-Hut batch: the same, on its own stream --------------------------
            if bh_inflight and self.bh.idle():
                if self.final_forces and any(rem <= bh_inflight for _, rem in bh_jobs.values()):
                    self.bh.step(1, _lib.BATCH_ACCELERATE)
                for sid in list(bh_jobs):
                    job, rem = bh_jobs[sid]
                    rem -= bh_inflight
                    job_steps += bh_inflight
                    body_steps += bh_inflight * job.n
                    if rem <= 0:
                        res = self.bh.download(sid)
                        self.bh.remove(sid)
                        del bh_jobs[sid]
                        finish(job, res)
                        if self.count_bh_interactions:
                            interactions += self._count(job, res)
                    else:
                        bh_jobs[sid] = (job, rem)
                bh_inflight = 0

            # -- admission: per-class queues, the budget granted in arrival order --------
            # A job whose class resource is busy keeps its place for the budget (the bytes
            # of such jobs ahead are held back) but does not stop later jobs of the other
            # class; under "fcfs" it stops every later job.
            blocked: set[int] = set()
            reserved = 0
            i = 0
            while i < len(waiting):
                job = waiting[i]
                if job.mode in blocked:
                    reserved += job.device_bytes()
                    i += 1
                    continue
                if in_use + reserved + job.device_bytes() > self.budget:
                    break
                slot = None
                if job.mode == _lib.MODE_N2:
                    busy = batch_inflight > 0       # joins at the next chunk boundary
                elif self.bh is not None:
                    busy = bh_inflight > 0 or len(bh_jobs) >= _lib.BH_BATCH_MAX_SYSTEMS
                else:
                    slot = next((s for s in self.slots if s.job is None), None)
                    busy = slot is None
                if busy:
                    if self.admission == "fcfs":
                        break
                    blocked.add(job.mode)
                    reserved += job.device_bytes()
                    i += 1
                    continue
                if job.mode == _lib.MODE_N2:
                    sid = (self.batch.add(job.bodies, job.params) if job.bodies is not None
                           else self.batch.add_generated(job.ics, job.n, job.params))
                    direct[sid] = (job, job.steps)
                    job.interactions_per_step = float(job.n) ** 2
                elif self.bh is not None:
                    job.params.leaf_size = self._leaf_for(job, None)
                    sid = (self.bh.add(job.bodies, job.params) if job.bodies is not None
                           else self.bh.add_generated(job.ics, job.n, job.params))
                    bh_jobs[sid] = (job, job.steps)
                else:
                    self._load(slot.ctx, job)
                    job.params.leaf_size = self._leaf_for(job, slot.ctx)
                    slot.job, slot.remaining, slot.inflight = job, job.steps, 0
                del waiting[i]
                job.admitted = now()
                in_use += job.device_bytes()
                peak = max(peak, in_use)

            if direct and not batch_inflight:
                k = min(self.chunk, min(rem for _, rem in direct.values()))
                self.batch.step(k)
                batch_inflight = k
            if bh_jobs and not bh_inflight:
                k = min(self.chunk, min(rem for _, rem in bh_jobs.values()))
                self.bh.step(k, _lib.BATCH_GRAPH if self.bh_graph else 0)
                bh_inflight = k

            # -- Barnes-Hut slots ---------------------------------------------------------
            for s in self.slots:
                if s.job is None:
                    continue
                if s.inflight and s.ctx.idle():
                    s.remaining -= s.inflight
                    job_steps += s.inflight
                    body_steps += s.inflight * s.job.n
                    s.inflight = 0
                    if s.remaining <= 0:
                        if self.final_forces:
                            s.ctx.accelerate(s.job.params)
                        res = s.ctx.download_bodies()
                        job, s.job = s.job, None
                        finish(job, res)
                        if self.count_bh_interactions:
                            interactions += self._count(job, res)   # at the final positions
                        continue
                if not s.inflight:
                    k = min(self.chunk, s.remaining)
                    if self.bh_graph:
                        s.ctx.step_graph(s.job.params, k)
                    else:
                        s.ctx.step(s.job.params, k)
                    s.inflight = k

            max_conc = max(max_conc, len(direct) + len(bh_jobs) + sum(1 for s in self.slots if s.job is not None))
            if not waiting and pending and not direct and not bh_jobs and all(s.job is None for s in self.slots):
                # nothing on the device: wait for the next arrival
                time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
            else:
                time.sleep(20e-6)

        wall = now()
        lat = np.array([j.latency for j in finished])
        return Report(len(finished), wall, job_steps, body_steps, interactions, lat, peak, max_conc)


def run_sequential(jobs: list[Job], clock=time.perf_counter) -> Report:
    """The same jobs one at a time, each alone on one context through the plain
    single-system path (nbhip_step: one force launch per direct-sum step with that path's
    own launch shape; the Barnes-Hut build and walk per step), in arrival order, every job
    queued at the start.  The baseline the continuous-batching server is compared with."""
    t0 = clock()
    lat = []
    job_steps = 0
    body_steps = 0.0
    with _lib.Context(0) as ctx:
        for j in sorted(jobs, key=lambda j: j.arrival):
            if j.bodies is not None:
                ctx.upload_bodies(j.bodies)
            else:
                ctx.generate(j.ics, j.n)
            ctx.step(j.params, j.steps)
            ctx.download_bodies()
            job_steps += j.steps
            body_steps += j.steps * j.n
            lat.append(clock() - t0)
    wall = clock() - t0
    return Report(len(jobs), wall, job_steps, body_steps, 0.0, np.array(lat), 0, 1)


def make_jobs(count: int, seed: int, n_range=(1024, 262144), steps_range=(10, 1000),
              direct_below: int = 16384, eps: float = 0.01, dt: float = 1e-3,
              theta: float = 0.5, leaf_size: int = 16, size: float = 128.0,
              device_ics: bool = False, bh_auto_root: bool = True) -> list[Job]:
    """The served workload: Plummer or disk+bulge systems, N log-uniform, steps uniform;
    direct sum below `direct_below` bodies, Barnes-Hut above.  Bodies are host arrays
    (numpy's generators, nbody_amd.ics) or, with device_ics, generator specs drawn on the
    device at admission.  Arrival times are zero; set them with poisson_arrivals()."""
    from . import ics
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        disk = bool(rng.integers(0, 2))
        job_seed = seed * 7919 + k
        spec = None
        if device_ics:
            spec = _lib.ics(_lib.IC_DISK_BULGE if disk else _lib.IC_PLUMMER, seed=job_seed)
            bodies = None
        else:
            bodies = ics.disk_bulge(n, seed=job_seed) if disk else ics.plummer(n, seed=job_seed)
        if n < direct_below:
            p = _lib.params(dt=dt, eps=eps, size=size, wrap=1)
        else:
            # the octree's root cube is found from the job's own bodies each step (size 0: no
            # torus wrap), so a system that spreads out is never clamped into edge octants
            p = _lib.params(mode=_lib.MODE_NLOGN, dt=dt, eps=eps, size=0.0 if bh_auto_root else size,
                            wrap=1, theta=theta, leaf_size=leaf_size)
        out.append(Job(bodies=bodies, steps=steps, params=p, tag="disk" if disk else "plummer",
                       ics=spec, n_bodies=n))
    return out


def poisson_arrivals(jobs: list[Job], rate: float, seed: int) -> None:
    """Arrival times of a Poisson stream of `rate` jobs per second, in list order."""
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, len(jobs))
    t = np.cumsum(gaps) - gaps[0]
    for j, a in zip(jobs, t):
        j.arrival = float(a)
