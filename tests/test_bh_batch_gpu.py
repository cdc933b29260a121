"""Barnes-Hut batches (nbhip_bh_batch_*): many systems, one build and one walk per step.

Bars:
  * a system's result is bit-identical alone in a batch and in any mix, eager or replayed
    from a graph, across joins and removals of other systems;
  * against one context running the same system with the walk unsplit (split 1), bit for
    bit wherever no two of its bodies share an 18-level cell (checked from the keys before
    each compared step): the batch's build and walk are the single-system ones;
  * against the oracle's bucket walk cut at 18 levels (orc_bh_accelerate_buckets with
    max_level 18, pinned on the CPU against the bucket tree's definition cut at 18 levels):
    >= 99.99% identical per-body visit counts (99.9% for groups of bodies packed into
    18-level cells), those bodies within 1e-5 (99.9%) / 1e-4 (all).
"""
import numpy as np
import pytest

from nbody_amd import _lib, ics

pytestmark = pytest.mark.gpu


def bh(**kw):
    kw.setdefault("mode", _lib.MODE_NLOGN)
    kw.setdefault("eps", 0.01)
    kw.setdefault("dt", 1e-3)
    kw.setdefault("theta", 0.5)
    return _lib.params(**kw)


def close_groups(n, size, seed, k=6):
    # groups of k bodies 2^-21 .. 2^-18 root edges from the first of the group
    rng = np.random.default_rng(seed)
    b = ics.plummer(n, seed=seed)
    lead = (np.arange(n) // k) * k
    d = rng.normal(size=(n, 3)) * (size * 2.0 ** rng.uniform(-21, -18, (n, 1)))
    d[lead == np.arange(n)] = 0.0
    b[:, 0:3] = (b[lead, 0:3] + d).astype(np.float32)
    return b


def systems():
    """(name, bodies, params): sizes 1 .. 20000, fixed and found roots, the cutoff law,
    both opening criteria, bucket sizes 1 .. 16, a system no larger than its bucket."""
    cut = ics.plummer(3000, seed=5)
    cut[::3, 3] = 2e-3   # radii: the reference cutoff law
    return [
        ("plummer_auto", ics.plummer(5000, seed=1), bh(size=0.0, leaf_size=16)),
        ("disk_wrap", ics.disk_bulge(20000, seed=2), bh(size=128.0, leaf_size=8, wrap=1)),
        ("one", ics.plummer(1, seed=3), bh(size=64.0)),
        ("two", ics.plummer(2, seed=4), bh(size=64.0, leaf_size=1)),
        ("bucket_root", ics.plummer(17, seed=6), bh(size=64.0, leaf_size=32)),
        ("cutoff", cut, bh(size=64.0, eps=0.0, leaf_size=4)),
        ("textbook", ics.disk_bulge(4000, seed=7), bh(size=128.0, theta=0.6, leaf_size=4,
                                                      opening=_lib.OPEN_TEXTBOOK)),
        ("groups", close_groups(3000, 64.0, 8), bh(size=64.0, leaf_size=1)),
    ]


def run(batch, members, steps):
    """Add members (name, bodies, params) in order, take the step plan, download each."""
    ids = {name: batch.add(b, p) for name, b, p in members}
    for k, flags in steps:
        batch.step(k, flags)
    out = {name: batch.download(i) for name, i in ids.items()}
    for i in ids.values():
        batch.remove(i)
    return out


PLAN = [(2, 0), (3, _lib.BATCH_GRAPH), (3, _lib.BATCH_GRAPH), (1, _lib.BATCH_ACCELERATE)]


def test_alone_and_mixed_are_identical(ctx):
    sys_ = systems()
    with _lib.BhBatch(ctx) as batch:
        alone = {}
        for m in sys_:
            alone.update(run(batch, [m], PLAN))
        rng = np.random.default_rng(0)
        for order in (list(range(len(sys_))), list(rng.permutation(len(sys_)))):
            mixed = run(batch, [sys_[i] for i in order], PLAN)
            for name, _, _ in sys_:
                assert np.array_equal(alone[name], mixed[name]), name
    # something moved and every force is finite
    for name, b, _ in sys_:
        assert np.all(np.isfinite(alone[name])), name
        if b.shape[0] > 1:
            assert not np.array_equal(alone[name][:, 0:3], b[:, 0:3]), name


def test_removals_and_joins_between_steps(ctx):
    sys_ = systems()
    with _lib.BhBatch(ctx) as ref, _lib.BhBatch(ctx) as batch:
        alone = {name: run(ref, [(name, b, p)], [(4, 0), (1, _lib.BATCH_ACCELERATE)])[name] for name, b, p in sys_}
        ids = {name: batch.add(b, p) for name, b, p in sys_[:5]}
        batch.step(2, _lib.BATCH_GRAPH)
        batch.remove(ids.pop("disk_wrap"))   # the systems after it move down
        batch.remove(ids.pop("one"))
        for name, b, p in sys_[5:]:
            ids[name] = batch.add(b, p)      # and join at the end, two steps behind
        batch.step(2, _lib.BATCH_GRAPH)
        head = {name: batch.download(ids[name]) for name in ("plummer_auto", "two", "bucket_root")}
        batch.step(1, _lib.BATCH_ACCELERATE)
        for name in ("plummer_auto", "two", "bucket_root"):
            assert np.array_equal(batch.download(ids[name])[:, 0:8], head[name][:, 0:8]), name
            assert np.array_equal(batch.download(ids[name]), alone[name]), name
        assert batch.info().systems == 6


def test_graph_replays_equal_eager_steps(ctx):
    sys_ = systems()[:4]
    with _lib.BhBatch(ctx) as a, _lib.BhBatch(ctx) as g:
        ia = [a.add(b, p) for _, b, p in sys_]
        ig = [g.add(b, p) for _, b, p in sys_]
        for _ in range(3):
            a.step(2)
            g.step(2, _lib.BATCH_GRAPH)
            # eager work on the graph's stream between replays
            g.step(1, _lib.BATCH_ACCELERATE)
            a.step(1, _lib.BATCH_ACCELERATE)
            for x, y in zip(ia, ig):
                assert np.array_equal(a.download(x), g.download(y))


def test_graph_chunks_of_any_length_equal_eager(ctx):
    # one step is captured and replayed once per step of a chunk: chunks of 3, 1 and 4
    # steps reuse it, a removal (the systems after the removed one move down) captures it
    # again; every chunk equals eager stepping bit for bit
    sys_ = systems()[:5]
    with _lib.BhBatch(ctx) as a, _lib.BhBatch(ctx) as g:
        ia = {name: a.add(b, p) for name, b, p in sys_}
        ig = {name: g.add(b, p) for name, b, p in sys_}

        def same():
            for name in ia:
                assert np.array_equal(a.download(ia[name]), g.download(ig[name])), name

        for k in (3, 1, 4):
            a.step(k)
            g.step(k, _lib.BATCH_GRAPH)
            same()
        a.remove(ia.pop("disk_wrap"))
        g.remove(ig.pop("disk_wrap"))
        for k in (2, 5):
            a.step(k)
            g.step(k, _lib.BATCH_GRAPH)
            same()


def keys18_unique(ctx, p):
    keys, _ = ctx.bh_keys(p)
    k18 = keys >> np.uint64(9)
    return np.unique(k18).size == k18.size


@pytest.mark.parametrize("name", ["plummer_auto", "disk_wrap", "two", "bucket_root", "cutoff", "textbook"])
def test_matches_one_context_unsplit(ctx, name):
    sys_ = {s[0]: s for s in systems()}
    _, b, p = sys_[name]
    others = [s for s in systems() if s[0] in ("groups", "one")]
    with _lib.BhBatch(ctx) as batch:
        for m in others[:1]:
            batch.add(m[1], m[2])
        sid = batch.add(b, p)
        for m in others[1:]:
            batch.add(m[1], m[2])
        ctx.upload_bodies(b)
        ctx.bh_configure(walk_split=1)
        try:
            compared = 0
            for step in range(4):
                unique = keys18_unique(ctx, p)
                ctx.step(p, 1)
                batch.step(1)
                got, want = batch.download(sid), ctx.download_bodies()
                if unique:
                    assert np.array_equal(got, want), (name, step)
                    compared += 1
                else:
                    ctx.upload_bodies(got)   # follow the batch's state
            assert compared >= 3, name
        finally:
            ctx.bh_configure(walk_split=0)


def rel_err(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


@pytest.mark.parametrize("leaf", [1, 16])
def test_forces_match_18_level_bucket_oracle(ctx, orc, leaf):
    # groups: an aggregate of a group applied within the softening length moves a force by
    # G M delta / eps^3 for the FP32 rounding delta of its centre (1e-6 for a 48-body group
    # at eps 0.1, 3e-5 of the force it lands on), so they run at eps 0.3, as the clumps of
    # the single-system bucket test do
    members = [("groups", close_groups(6000, 64.0, 11 + leaf, k=6 if leaf < 4 else 48), 64.0, 0.3, 0.5),
               ("disk", ics.disk_bulge(20000, seed=12 + leaf), 128.0, 0.01, 0.5),
               ("plummer", ics.plummer(8000, seed=13 + leaf), 64.0, 0.01, 0.7)]
    with _lib.BhBatch(ctx) as batch:
        ids = [batch.add(b, bh(size=size, eps=eps, theta=theta, leaf_size=leaf))
               for _, b, size, eps, theta in members]
        batch.step(1, _lib.BATCH_ACCELERATE | _lib.BATCH_COUNT)
        for (name, b, size, eps, theta), sid in zip(members, ids):
            out = batch.download(sid)
            visits = batch.visits(sid)
            x = b.copy()
            _, xv = orc.bh_accelerate_buckets(x, size, theta=theta, leaf_size=leaf, eps=eps, max_level=18)
            eq = visits == xv
            e = rel_err(out[:, 8:11], x[:, 8:11])
            print(f"{name} leaf {leaf}: same traversal {eq.mean():.5f}, 99.9% {np.quantile(e[eq], 0.999):.2e} "
                  f"max {e[eq].max():.2e}")
            assert eq.mean() >= (0.999 if name == "groups" else 0.9999), name
            assert np.quantile(e[eq], 0.999) <= 1e-5 and e[eq].max() <= 1e-4, name
            assert np.array_equal(out[:, 0:8], b[:, 0:8])   # accelerations only


def test_generated_systems_equal_uploaded(ctx):
    spec = _lib.ics(_lib.IC_DISK_BULGE, seed=77)
    n = 12345
    ctx.generate(spec, n)
    b = ctx.download_bodies()
    p = bh(size=0.0, leaf_size=16)
    with _lib.BhBatch(ctx) as batch:
        i = batch.add_generated(spec, n, p)
        j = batch.add(b, p)
        assert np.array_equal(batch.download(i), b)
        batch.step(3, _lib.BATCH_GRAPH)
        assert np.array_equal(batch.download(i), batch.download(j))


def test_many_one_body_systems_and_limits(ctx):
    with _lib.BhBatch(ctx) as batch:
        rng = np.random.default_rng(3)
        bodies = []
        for k in range(_lib.BH_BATCH_MAX_SYSTEMS):
            b = np.zeros((1, 12), np.float32)
            b[0, 0:3] = rng.normal(size=3)
            b[0, 4:7] = rng.normal(size=3)
            b[0, 7] = 1.0
            bodies.append(b)
            batch.add(b, bh(size=0.0 if k % 2 else 16.0, wrap=0))
        with pytest.raises(_lib.HipError):
            batch.add(bodies[0], bh(size=16.0))
        batch.step(2, _lib.BATCH_GRAPH)
        info = batch.info()
        assert info.systems == _lib.BH_BATCH_MAX_SYSTEMS and info.bodies == _lib.BH_BATCH_MAX_SYSTEMS
        assert info.levels == _lib.BH_BATCH_LEVELS
        for sid in (1, 2, 256, 512):
            out = batch.download(sid)
            b = bodies[sid - 1]
            assert np.all(out[0, 8:11] == 0.0)   # a body alone feels nothing
            x = b[0, 0:3].astype(np.float32)
            v = b[0, 4:7].astype(np.float32)
            for _ in range(2):
                x = (x + v * np.float32(1e-3)).astype(np.float32)
            assert np.array_equal(out[0, 0:3], x)


def test_partial_sorts_leave_the_same_bodies(ctx, monkeypatch):
    # The batch's one sort over every system's keys may skip low-digit passes (fixing the
    # runs they leave): every fixed skip and the adaptive choice give the full sort's steps,
    # bit for bit, across a join (which makes the next sort a full one) and a removal (after
    # which the last sort's choice stands: the remaining keys keep their order)
    members = systems()
    res = {}
    for skip in ("0", "2", "4", ""):
        if skip:
            monkeypatch.setenv("NBHIP_RS_SKIP", skip)
        else:
            monkeypatch.delenv("NBHIP_RS_SKIP", raising=False)
        with _lib.BhBatch(ctx) as batch:
            ids = {name: batch.add(b, p) for name, b, p in members[:5]}
            batch.step(3)
            batch.step(3, _lib.BATCH_GRAPH)
            name, b, p = members[5]
            ids[name] = batch.add(b, p)
            batch.step(2, _lib.BATCH_GRAPH)
            batch.remove(ids.pop("disk_wrap"))
            batch.step(2)
            res[skip] = {name: batch.download(i) for name, i in ids.items()}
    for k, v in res.items():
        for name, out in v.items():
            assert np.array_equal(out.view(np.uint32), res["0"][name].view(np.uint32)), (k, name)


def test_errors(ctx):
    b = ics.plummer(100, seed=1)
    with _lib.BhBatch(ctx) as batch:
        batch.step(3)                      # an empty batch: nothing to do
        with pytest.raises(_lib.HipError):
            batch.add(b, _lib.params(mode=_lib.MODE_N2))
        # the parameter checks of a single context's Barnes-Hut step
        for bad in (bh(eps=-1.0), bh(leaf_size=0), bh(size=-1.0), bh(theta=-0.5), bh(theta=float("nan")),
                    bh(opening=7)):
            with pytest.raises(_lib.HipError):
                batch.add(b, bad)
            with pytest.raises(_lib.HipError):
                ctx.upload_bodies(b)
                ctx.step(bad, 1)
        assert batch.info().systems == 0 and batch.info().bodies == 0
        i = batch.add(b, bh(size=64.0))
        with pytest.raises(_lib.HipError):
            batch.step(1, _lib.BATCH_COUNT)    # counting needs accelerations only
        with pytest.raises(_lib.HipError):
            batch.step(1, 64)
        with pytest.raises(_lib.HipError):
            batch.visits(i)                    # no counting step yet
        with pytest.raises(_lib.HipError):
            batch.remove(i + 1)
        batch.step(0)
        batch.step(1, _lib.BATCH_ACCELERATE | _lib.BATCH_COUNT)
        assert batch.visits(i).sum() > 0
        batch.remove(i)
        assert batch.info().systems == 0 and batch.info().bodies == 0
        batch.step(2)


def test_profiled_steps_equal_plain_steps(ctx):
    sys_ = systems()[:3]
    with _lib.BhBatch(ctx) as a, _lib.BhBatch(ctx) as b:
        ia = [a.add(x, p) for _, x, p in sys_]
        ib = [b.add(x, p) for _, x, p in sys_]
        build_ms, walk_ms = a.profile(3)
        b.step(3)
        assert build_ms > 0.0 and walk_ms > 0.0
        for x, y in zip(ia, ib):
            assert np.array_equal(a.download(x), b.download(y))
