"""The batched direct-sum path (many independent systems per launch), through the C ABI.

Bars:
  * a system stepped in a batch is bit-identical to the same system stepped alone by the
    single-system path with the same tile shape and no j-split (the batch runs each
    body's whole j loop in one block), and bit-identical whatever else is in the batch,
    in persistent or grid mode, eagerly or replayed from a hipGraph;
  * against the unbatched per-system path with its own launch shape: accelerations to a
    relative 2e-6 (only the grouping of the compensated partial sums differs);
  * against the FP64 direct sum: 1e-5 per body.
"""
import numpy as np
import pytest

from nbody_amd import _lib, ics

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def rel_err(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


def systems():
    out = []
    for k, n in enumerate([1, 65, 700, 1000, 2048, 3001, 300]):
        b = ics.plummer(n, seed=500 + k)
        p = _lib.params(dt=1e-3 * (1 + k % 3), eps=[0.01, 0.0, 0.02][k % 3], G=[1.0, 2.0][k % 2],
                        size=[10000.0, 8.0][k % 2], wrap=1)
        if k == 4:
            b[:, 3] = 0.01     # radius cutoffs
        out.append((b, p))
    return out


def alone(ctx, b, p, steps, monkeypatch, P=2, W=8):
    monkeypatch.setenv("NBHIP_DIRECT_P", str(P))
    monkeypatch.setenv("NBHIP_DIRECT_W", str(W))
    monkeypatch.setenv("NBHIP_DIRECT_JSPLIT", "1")
    ctx.upload_bodies(b)
    ctx.step(p, steps)
    return ctx.download_bodies()


@pytest.mark.parametrize("persistent", [True, False])
def test_batch_equals_alone_bit_exact(ctx, monkeypatch, persistent):
    sy = systems()
    with _lib.Batch(ctx) as batch:
        batch.configure(2, 8, persistent)
        ids = [batch.add(b, p) for b, p in sy]
        batch.step(3)
        got = [batch.download(i) for i in ids]
    for (b, p), g in zip(sy, got):
        ref = alone(ctx, b, p, 3, monkeypatch)
        assert np.array_equal(bits(g), bits(ref)), b.shape[0]


def test_same_alone_and_in_the_mix(ctx):
    target = ics.plummer(1500, seed=77)
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0)
    with _lib.Batch(ctx) as a:
        i = a.add(target, p)
        a.step(4)
        solo = a.download(i)
    with _lib.Batch(ctx) as m:
        rng = np.random.default_rng(3)
        for k in range(20):
            n = int(rng.integers(100, 5000))
            m.add(ics.plummer(n, seed=900 + k), _lib.params(dt=2e-3, eps=0.02))
        j = m.add(target, p)
        for k in range(5):
            m.add(ics.plummer(300 + k, seed=1900 + k), _lib.params(dt=1e-3, eps=0.0))
        m.step(4)
        mixed = m.download(j)
    assert np.array_equal(bits(solo), bits(mixed))


def test_graph_replay_equals_eager(ctx):
    sy = systems()
    res = []
    for flags in [0, _lib.BATCH_GRAPH]:
        with _lib.Batch(ctx) as batch:
            ids = [batch.add(b, p) for b, p in sy]
            for _ in range(3):                 # replay the captured graph several times,
                batch.step(4, flags)            # with eager work and syncs in between
                _ = batch.download(ids[2])
                ctx.synchronize()
            res.append([batch.download(i) for i in ids])
    for a, g in zip(*res):
        assert np.array_equal(bits(a), bits(g))


def test_odd_steps_and_late_joiners(ctx, monkeypatch):
    # a system added after an odd number of steps starts from its own state
    sy = systems()
    with _lib.Batch(ctx) as batch:
        first = batch.add(*sy[3])
        batch.step(3, _lib.BATCH_GRAPH)
        late = batch.add(*sy[5])
        batch.step(2)
        batch.remove(first)
        batch.step(1, _lib.BATCH_GRAPH)
        got = batch.download(late)
    ref = alone(ctx, sy[5][0], sy[5][1], 3, monkeypatch)
    assert np.array_equal(bits(got), bits(ref))


def test_per_system_path_agrees(ctx):
    sy = systems()
    out = []
    for flags in [0, _lib.BATCH_PER_SYSTEM]:
        with _lib.Batch(ctx) as batch:
            ids = [batch.add(b, p) for b, p in sy]
            batch.step(1, flags | _lib.BATCH_ACCELERATE)
            out.append([batch.download(i) for i in ids])
    for a, b in zip(*out):
        if a.shape[0] > 1:
            assert rel_err(a[:, 8:11], b[:, 8:11]).max() <= 2e-6


def test_tuned_batch_matches_f64(ctx, orc):
    rng = np.random.default_rng(8)
    sy = [(ics.plummer(int(n), seed=k), _lib.params(dt=1e-3, eps=0.01))
          for k, n in enumerate(np.exp(rng.uniform(np.log(256), np.log(8192), 24)).astype(int))]
    with _lib.Batch(ctx) as batch:
        ids = [batch.add(b, p) for b, p in sy]
        ms = batch.tune()
        info = batch.info()
        print(f"tuned: P={info.P} W={info.W} launch={info.launch} {ms:.3f} ms/step")
        assert info.systems == len(sy) and info.interactions_per_step == sum(float(b.shape[0]) ** 2 for b, _ in sy)
        batch.step(1, _lib.BATCH_ACCELERATE)
        for (b, _), i in zip(sy[:6], ids[:6]):
            a = batch.download(i)[:, 8:11]
            assert rel_err(a, orc.direct_f64(b, eps=0.01)).max() <= 1e-5


def test_rejects_bad_input(ctx):
    with _lib.Batch(ctx) as batch:
        with pytest.raises(_lib.HipError):
            batch.add(ics.plummer(10, seed=1), _lib.params(mode=_lib.MODE_NLOGN))
        with pytest.raises(_lib.HipError):
            batch.remove(12345)
        with pytest.raises(_lib.HipError):
            batch.configure(3, 3)


def test_graph_after_long_eager_call_equals_eager(ctx):
    # an eager call of more steps than the step counters hold reallocates them; a graph
    # captured before must not be replayed on the freed counters afterwards
    sy = systems()[:4]
    res = []
    for graph in (True, False):
        with _lib.Batch(ctx) as batch:
            ids = [batch.add(b, p) for b, p in sy]
            f = _lib.BATCH_GRAPH if graph else 0
            batch.step(2, f)
            batch.step(300)
            batch.step(2, f)
            res.append([batch.download(i) for i in ids])
    for g, e in zip(*res):
        assert np.array_equal(bits(g), bits(e))


def test_rejects_systems_past_the_body_limit(ctx):
    spec = _lib.ics(_lib.IC_PLUMMER)
    with _lib.Batch(ctx) as batch:
        with pytest.raises(_lib.HipError) as e:
            batch.add_generated(spec, _lib.MAX_BODIES + 1, _lib.params(eps=0.01))
        assert e.value.status == _lib.ERR_INVALID


def c4_workload():
    # the benchmark's 512 systems: N log-uniform in [256, 32768], seeds 2026 + k
    rng = np.random.default_rng(2026)
    ns = np.exp(rng.uniform(np.log(256), np.log(32768), 512)).astype(int)
    return [ics.plummer(int(n), seed=2026 + k) for k, n in enumerate(ns)]


def test_c4_workload_50_step_graph(ctx, orc):
    # The full C4 workload as the benchmark runs it: autotuned, 50 fused steps replayed from
    # one hipGraph; then one accelerate-only pass, whose forces on sampled bodies of the 8
    # largest and 8 smallest systems must be within 1e-5 of FP64.  Three systems (one above
    # 16k bodies) stepped alone in a batch of one with the same launch shape must match
    # their state in the full batch bit for bit.
    systems = c4_workload()
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0, wrap=1)
    with _lib.Batch(ctx) as batch:
        ids = [batch.add(b, p) for b in systems]
        batch.tune()
        info = batch.info()
        assert info.systems == 512
        batch.step(50, _lib.BATCH_GRAPH)
        batch.step(1, _lib.BATCH_ACCELERATE)
        order = np.argsort([b.shape[0] for b in systems])
        check = list(order[:8]) + list(order[-8:])
        big = int(next(k for k in order[::-1] if 16384 < systems[k].shape[0]))
        alone_pick = [int(order[0]), int(order[len(order) // 2]), big]
        got = {k: batch.download(ids[k]) for k in set(check) | set(alone_pick)}
    rng = np.random.default_rng(4)
    for k in check:
        g = got[k]
        s = rng.choice(g.shape[0], min(g.shape[0], 128), replace=False).astype(np.uint32)
        a64 = orc.direct_f64(g, s, eps=0.01)
        e = rel_err(g[s, 8:11], a64)
        assert e.max() <= 1e-5, (k, g.shape[0], e.max())
    for k in alone_pick:
        with _lib.Batch(ctx) as one:
            one.configure(info.P, info.W, info.launch)
            i = one.add(systems[k], p)
            one.step(50, _lib.BATCH_GRAPH)
            one.step(1, _lib.BATCH_ACCELERATE)
            solo = one.download(i)
        assert np.array_equal(bits(solo), bits(got[k])), (k, systems[k].shape[0])


@pytest.mark.parametrize("P", [2, 4, 8])
def test_pairs_mode_equals_symmetric_pass_alone(ctx, monkeypatch, P):
    # pairs mode runs each system through the pair-symmetric tile with blocks of 64*P*4:
    # bit-identical to the single-system pair-symmetric pass of the same shape, fused steps
    # and radius cutoffs included, and the same alone and in a mix
    sy = systems() + [(ics.plummer(9000, seed=77), _lib.params(dt=1e-3, eps=0.01))]
    with _lib.Batch(ctx) as batch:
        batch.configure(P, 4, _lib.BATCH_LAUNCH_PAIRS)
        ids = [batch.add(b, p) for b, p in sy]
        assert batch.info().launch == _lib.BATCH_LAUNCH_PAIRS
        batch.step(3)
        got = [batch.download(i) for i in ids]
    monkeypatch.setenv("NBHIP_DIRECT_SYM", "1")
    monkeypatch.setenv("NBHIP_DIRECT_P", str(P))
    monkeypatch.setenv("NBHIP_DIRECT_W", "4")
    for (b, p), g in zip(sy, got):
        ctx.upload_bodies(b)
        ctx.step(p, 3)
        ref = ctx.download_bodies()
        assert np.array_equal(bits(g), bits(ref)), b.shape[0]
    with _lib.Batch(ctx) as one:
        one.configure(P, 4, _lib.BATCH_LAUNCH_PAIRS)
        i = one.add(*sy[-1])
        one.step(3)
        assert np.array_equal(bits(one.download(i)), bits(got[-1]))


def test_pairs_mode_graph_and_f64(ctx, orc):
    sy = systems()[1:]
    res = []
    for flags in (0, _lib.BATCH_GRAPH):
        with _lib.Batch(ctx) as batch:
            batch.configure(4, 4, _lib.BATCH_LAUNCH_PAIRS)
            ids = [batch.add(b, p) for b, p in sy]
            for _ in range(3):
                batch.step(4, flags)
                _ = batch.download(ids[1])
                ctx.synchronize()
            batch.step(1, _lib.BATCH_ACCELERATE)
            res.append([batch.download(i) for i in ids])
    for a, g in zip(*res):
        assert np.array_equal(bits(a), bits(g))
    for (b, p), g in zip(sy, res[0]):
        a64 = orc.direct_f64(g, G=float(p.G), eps=float(p.eps))
        assert rel_err(g[:, 8:11], a64).max() <= 1e-5


def test_rejects_bad_launch_modes(ctx):
    with _lib.Batch(ctx) as batch:
        with pytest.raises(_lib.HipError):
            batch.configure(2, 4, 3)
        with pytest.raises(_lib.HipError):
            batch.configure(1, 8, _lib.BATCH_LAUNCH_PAIRS)   # no pairs kernel for this shape
