"""Parity of the HIP Barnes-Hut path with the reference octree, through the C ABI.

Bars (stated per check):
  * Morton keys and the sort permutation: bit-exact against the oracle's restatement of
    the reference's octant descent and std::stable_sort (and against keys written from
    the reference's own Bounds code, tests/golden/cell_keys.npz);
  * forces with leaf size 1 and the reference opening criterion at the same theta:
      - against the oracle's FP64-centre Barnes-Hut (the reference tree and traversal
        with exact centres of mass): >= 99.99% identical traversals, and those bodies
        within a relative 1e-5 (99.9%) / 1e-4 (all);
      - against the reference itself: >= 99.9% identical traversals (same number of
        applied nodes), and for those bodies |a_gpu - a_ref| / |a_ref| <= (the
        reference's own distance from the FP64-centre oracle) + 1e-4.  The reference folds
        FP32 means in insertion order (source/bhtree.cpp:87-95); that rounding alone moves
        its forces by ~2e-4 at 2e4 bodies and ~1e-3 at 2^20, and a node lying within it
        of a body's acceptance threshold flips between opened and applied, which with the
        reference's criterion moves that body's force by up to ~10% -- counted and
        bounded, not hidden;
  * forces with leaf size > 1 (buckets): against the oracle's bucket walk over the
    reference tree with FP64 centres (oracle orc_bh_accelerate_buckets, pinned on the CPU
    against the leaf-1 oracle, the FP64 direct sum and the bucket tree's definition):
    >= 99.99% identical traversals, those bodies within 1e-5 (99.9%) / 1e-4 (all);
  * theta large enough that every cell opens: equal to the direct sum to 1e-5.
"""
import os

import numpy as np
import pytest

from bh_definition import bucket_tree, walk
from nbody_amd import _lib, ics

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_err(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


def bh_params(**kw):
    kw.setdefault("mode", _lib.MODE_NLOGN)
    return _lib.params(**kw)


@pytest.mark.parametrize("size", [256, 10000, 100])
def test_keys_match_reference_golden(ctx, size):
    g = np.load(os.path.join(GOLD, "cell_keys.npz"))
    ctx.upload_bodies(g["bodies"])
    keys, perm = ctx.bh_keys(bh_params(size=float(size)))
    assert np.array_equal(keys, g["keys_%d" % size])
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))


# 2^19 - 1 and 2^19 keys sit either side of the switch from 2048- to 4096-key radix tiles
@pytest.mark.parametrize("n,size", [(1, 64.0), (2, 64.0), (5000, 64.0), (100000, 128.0), (70001, 10000.0),
                                    ((1 << 19) - 1, 128.0), (1 << 19, 128.0)])
def test_keys_and_permutation_bit_exact(ctx, orc, n, size):
    b = ics.disk_bulge(n, seed=n) if n > 2 else ics.plummer(n, seed=n)
    if n > 10:
        b[7, 0:3] = b[3, 0:3]          # an exact duplicate: a tie in the stable sort
        b[9, 0:3] = [1e6, -1e6, 0.0]   # a body outside the root cube
    ctx.upload_bodies(b)
    keys, perm = ctx.bh_keys(bh_params(size=size))
    ok = orc.morton_keys(b, size, 21)
    assert np.array_equal(keys, ok)
    assert np.array_equal(perm, orc.stable_sort(ok))


@pytest.mark.parametrize("skip", [0, 1, 2, 3, 4])
def test_partial_sorts_equal_the_full_sort(ctx, orc, monkeypatch, skip):
    # A sort may skip up to four low-digit passes; k_rs_fix then sorts the runs of equal key
    # prefixes those leave in input order, by (key, input position).  Whatever the number
    # skipped, keys and permutation equal the oracle's stable sort bit for bit: short runs
    # (disk + bulge: insertion), and runs far beyond the cap (a dense cluster under a wide
    # root: the heapsort path), with ties.
    monkeypatch.setenv("NBHIP_RS_SKIP", str(skip))
    cluster = ics.plummer(30000, seed=6, a=0.02)
    cluster[11, 0:3] = cluster[4, 0:3]
    for name, b, size in (("disk", ics.disk_bulge(100000, seed=5), 128.0), ("cluster", cluster, 10000.0)):
        ctx.upload_bodies(b)
        keys, perm = ctx.bh_keys(bh_params(size=size))
        ok = orc.morton_keys(b, size, 21)
        assert np.array_equal(keys, ok), name
        assert np.array_equal(perm, orc.stable_sort(ok)), name


def test_partial_sort_steps_equal_full_sort_steps(monkeypatch):
    # Steps with every sort a full one (NBHIP_RS_SKIP=0), with each fixed skip, and with the
    # skip chosen from the last sort's runs (unset) leave the same bodies, bit for bit,
    # eagerly and from graphs; so does a cluster whose runs exceed the cap, and a switch of
    # root size (all keys change) between steps.
    disk = ics.disk_bulge(60000, seed=9)
    cluster = ics.plummer(20000, seed=10, a=0.02)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, leaf_size=16, wrap=0)
    q = bh_params(size=10000.0, theta=0.5, eps=1e-3, dt=2e-3, leaf_size=1, wrap=0)
    res = {}
    for skip in ("0", "2", "3", "4", ""):
        if skip:
            monkeypatch.setenv("NBHIP_RS_SKIP", skip)
        else:
            monkeypatch.delenv("NBHIP_RS_SKIP", raising=False)
        out = []
        with _lib.Context(0) as c:
            c.upload_bodies(disk)
            c.step(p, 3)
            c.step_graph(p, 4)
            c.step_graph(p, 4)
            out.append(c.download_bodies())
            c.step(bh_params(size=256.0, theta=0.5, eps=0.01, dt=1e-3, leaf_size=16, wrap=0), 2)
            out.append(c.download_bodies())
            c.upload_bodies(cluster)
            c.step(q, 3)
            c.step_graph(q, 3)
            out.append(c.download_bodies())
        res[skip] = out
    for k, v in res.items():
        for a, ref in zip(v, res["0"]):
            assert np.array_equal(a.view(np.uint32), ref.view(np.uint32)), k


def test_adaptive_sort_skips_passes_on_key_ordered_steps(ctx, monkeypatch):
    # New bodies get a full sort; once a step has left them in key order with short runs
    # of equal key prefixes (disk + bulge under a 128 root: at most a few bodies share a
    # 40-bit prefix), the sorts skip low-digit passes; new bodies again: a full sort.
    monkeypatch.delenv("NBHIP_RS_SKIP", raising=False)
    b = ics.disk_bulge(1 << 18, seed=12)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, leaf_size=16, wrap=0)
    ctx.upload_bodies(b)
    ctx.step(p, 1)
    assert ctx.bh_sort_passes() == 8
    ctx.step(p, 2)
    assert ctx.bh_sort_passes() < 8
    ctx.step_graph(p, 3)
    assert ctx.bh_sort_passes() < 8
    ctx.upload_bodies(b)
    ctx.step(p, 1)
    assert ctx.bh_sort_passes() == 8


def check_tree(nodes, perm, b):
    """Structural invariants of the pre-order tree."""
    total = nodes.size
    assert total >= 1
    assert np.all(nodes["next"] > np.arange(total)) and np.all(nodes["next"] <= total)
    assert nodes["next"][0] == total
    # every body sits in exactly one leaf / bucket
    leaves = nodes[nodes["count"] > 0]
    cover = np.zeros(b.shape[0], int)
    for f, c in zip(leaves["first"], leaves["count"]):
        cover[f:f + c] += 1
    assert np.all(cover == 1)
    # the root holds all the mass; masses are sums over the subtree's leaves
    m_sorted = b[perm, 7].astype(np.float64)
    assert np.isclose(nodes["mass"][0], m_sorted.sum(), rtol=1e-6)
    for i in np.linspace(0, total - 1, 50).astype(int):
        sub = nodes[i:nodes["next"][i]]
        lv = sub[sub["count"] > 0]
        ms = sum(m_sorted[f:f + c].sum() for f, c in zip(lv["first"], lv["count"]))
        assert np.isclose(nodes["mass"][i], ms, rtol=1e-5)


@pytest.mark.parametrize("n,leaf", [(1, 1), (2, 1), (3, 1), (1000, 1), (20000, 1), (20000, 16)])
def test_tree_structure(ctx, n, leaf):
    b = ics.plummer(n, seed=3 * n + leaf)
    if n == 3:
        b[2, 0:3] = b[1, 0:3]   # coincident bodies share a 21-level cell
    ctx.upload_bodies(b)
    p = bh_params(size=64.0, leaf_size=leaf)
    nodes = ctx.bh_nodes(p)
    _, perm = ctx.bh_keys(p)
    check_tree(nodes, perm, b)
    if leaf > 1:
        assert nodes["count"].max() <= leaf or n <= leaf


@pytest.mark.parametrize("leaf", [1, 2, 5, 15, 16, 17, 40])
@pytest.mark.parametrize("kind", ["plummer", "clumps"])
def test_bucket_tree_matches_definition(ctx, leaf, kind):
    # the node array against the definition, at leaf sizes either side of the window
    # searches' 16-body limit; clumps put many bodies in a few deep cells and duplicates
    # in 21-level cells
    n = 6000
    if kind == "plummer":
        b = ics.plummer(n, seed=11 + leaf)
    else:
        rng = np.random.default_rng(leaf)
        b = ics.plummer(n, seed=leaf)
        centres = rng.normal(size=(12, 3)).astype(np.float32)
        pick = rng.integers(0, 12, n)
        b[:, 0:3] = centres[pick] + (rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-6, -1, (n, 1))).astype(np.float32)
        b[::37, 0:3] = b[1::37, 0:3][: b[::37].shape[0]]  # exact duplicates
    ctx.upload_bodies(b)
    p = bh_params(size=64.0, theta=0.7, leaf_size=leaf)
    nodes = ctx.bh_nodes(p)
    keys, perm = ctx.bh_keys(p)
    spm = b[perm][:, [0, 1, 2, 7]]
    ref = bucket_tree(np.sort(keys), spm, 64.0, 0.7, leaf)
    assert nodes.size == len(ref)
    r = np.array([row[:3] for row in ref], np.int64)
    assert np.array_equal(nodes["first"], r[:, 0])
    assert np.array_equal(nodes["count"], r[:, 1])
    assert np.array_equal(nodes["next"], r[:, 2])
    thr = np.array([row[3] for row in ref], np.float32)
    assert np.array_equal(nodes["open_sq"].view(np.uint32), thr.view(np.uint32))
    com = np.array([row[4:7] for row in ref], np.float64)
    mass = np.array([row[7] for row in ref], np.float64)
    assert np.allclose(nodes["mass"], mass, rtol=1e-6, atol=0)
    assert np.allclose(nodes["com"], com, rtol=1e-6, atol=1e-6)


def test_golden_reference_barnes_hut(ctx):
    # util::disk(512) with its radius cutoffs and central mass, root 10000 (the reference
    # default), theta 0.5 and 8: forces and per-body applied-node counts from the reference
    g = np.load(os.path.join(GOLD, "barnes_hut.npz"))
    for theta, tag in [(0.5, "05"), (8.0, "8")]:
        # leaves only: the golden disk has radius cutoffs and no softening
        ctx.upload_bodies(g["bodies"])
        p = bh_params(size=10000.0, theta=theta)
        visits, _ = ctx.bh_interactions(p)
        out = ctx.download_bodies()
        e = rel_err(out[:, 8:11], g["out_theta" + tag][:, 8:11])
        eq = visits == g["visits_theta" + tag]
        print(f"theta {theta}: max rel err {e.max():.2e}, visits equal {eq.mean():.4f}")
        assert eq.mean() >= 0.999
        assert e[eq].max() <= 1e-4


@pytest.mark.parametrize("n,size,theta", [(3000, 64.0, 0.5), (20000, 128.0, 0.5), (20000, 128.0, 2.0),
                                          (8000, 10000.0, 0.5)])
def test_forces_match_reference_barnes_hut(ctx, orc, n, size, theta):
    b = ics.disk_bulge(n, seed=n + 1)
    ctx.upload_bodies(b)
    p = bh_params(size=size, theta=theta, eps=0.0)
    visits, total = ctx.bh_interactions(p)
    out = ctx.download_bodies()
    nodes = orc.bh_build(b, size)
    x64 = b.copy()
    _, xv = orc.bh_accelerate_f64com(x64, size, theta=theta, eps=0.0, visits=True, nodes=nodes)
    e64 = rel_err(out[:, 8:11], x64[:, 8:11])
    eq64 = visits == xv
    ref = b.copy()
    _, rv = orc.bh_accelerate(ref, size, theta=theta, eps=0.0, visits=True, nodes=nodes)
    e = rel_err(out[:, 8:11], ref[:, 8:11])
    eq = visits == rv
    print(f"n={n} size={size} theta={theta}: vs FP64-centre BH: same traversal {eq64.mean():.5f}, "
          f"max {e64[eq64].max():.2e} 99.9% {np.quantile(e64[eq64], 0.999):.2e}; vs reference: same "
          f"traversal {eq.mean():.5f}, max {e[eq].max():.2e}, overall max {e.max():.2e}; "
          f"interactions {total}")
    assert eq64.mean() >= 0.9999
    assert np.quantile(e64[eq64], 0.999) <= 1e-5 and e64[eq64].max() <= 1e-4
    assert eq.mean() >= 0.999
    own = rel_err(x64[:, 8:11], ref[:, 8:11])   # the reference's own rounding
    assert np.all(e[eq] <= own[eq] + 1e-4)
    assert total == visits.sum()


def test_theta_zero_accepts_the_root(ctx):
    # Reference theta = 0 accepts a node whenever |com - x|^2 > 0, so every body sees the
    # root alone: a_i = G M (com - x_i) / (|com - x_i|^2 + eps^2)^{3/2}, its own mass included
    b = ics.plummer(3000, seed=12)
    ctx.upload_bodies(b)
    ctx.accelerate(bh_params(size=64.0, theta=0.0, eps=0.01, G=2.0))
    a = ctx.download_bodies()[:, 8:11].astype(np.float64)
    m = b[:, 7].astype(np.float64)
    x = b[:, 0:3].astype(np.float64)
    M = m.sum()
    com = (m[:, None] * x).sum(0) / M
    d = com - x
    want = 2.0 * M * d / ((d * d).sum(1) + 1e-4)[:, None] ** 1.5
    assert rel_err(a, want).max() <= 1e-5
    _, total = ctx.bh_interactions(bh_params(size=64.0, theta=0.0, eps=0.01))
    assert total == 3000


def test_coincident_bodies_match_fp64_centre_oracle(ctx, orc):
    # 600 bodies sharing one position: one 21-level cell, which the GPU keeps as a bucket
    # and the reference splits down to FLT_EPSILON; the coincident bodies open it, and
    # both add exactly zero for it (d = 0): the GPU body by body (599 more visits), the
    # reference through one merged leaf
    b = ics.plummer(3000, seed=13)
    b[:600, 0:3] = b[0, 0:3]
    ctx.upload_bodies(b)
    p = bh_params(size=64.0, theta=0.5, eps=0.01)
    visits, _ = ctx.bh_interactions(p)
    out = ctx.download_bodies()
    x64 = b.copy()
    _, xv = orc.bh_accelerate_f64com(x64, 64.0, theta=0.5, eps=0.01, visits=True)
    shared = np.zeros(len(b), bool)
    shared[:600] = True
    assert np.all(visits[shared] - xv[shared] == 599)
    same = visits == xv
    check = same | shared
    e = rel_err(out[check, 8:11], x64[check, 8:11])
    print(f"coincident: same traversal {same[~shared].mean():.5f}, max err {e.max():.2e}")
    assert same[~shared].mean() >= 0.999 and e.max() <= 1e-4


def test_bodies_outside_the_root_cube_match_reference(ctx, orc):
    # A Plummer sphere in a root cube of edge 8: 6% of bodies lie outside and go to edge
    # octants, as the reference's comparisons send them; a few beyond a corner in all
    # three coordinates share one 21-level key.  The reference merges such bodies into a
    # leaf at the first one's position (source/bhtree.cpp:37-42), which the FP64-centre
    # oracle (centres rebuilt from leaves) then spreads to every ancestor, so this case is
    # checked against the reference's own FP32 tree, whose ancestors accumulate the true
    # positions: identical traversals, forces within its FP32 centre rounding (a CPU
    # emulation of the GPU tree measured 7e-5 at 99.9%, 1.5e-4 max).
    b = ics.plummer(4000, seed=14)
    size = 8.0
    assert np.mean(np.abs(b[:, 0:3]).max(1) > size / 2) > 0.01
    ctx.upload_bodies(b)
    visits, _ = ctx.bh_interactions(bh_params(size=size, theta=0.5, eps=0.01))
    out = ctx.download_bodies()
    ref = b.copy()
    _, rv = orc.bh_accelerate(ref, size, theta=0.5, eps=0.01, visits=True)
    same = visits == rv
    e = rel_err(out[same, 8:11], ref[same, 8:11])
    print(f"outside: same traversal {same.mean():.5f}, 99.9% {np.quantile(e, 0.999):.2e}, max {e.max():.2e}")
    assert same.mean() >= 0.999
    assert np.quantile(e, 0.999) <= 2e-4 and e.max() <= 1e-3


def test_all_cells_open_equals_direct_sum(ctx, orc):
    b = ics.plummer(4000, seed=8)
    ctx.upload_bodies(b)
    ctx.accelerate(bh_params(size=64.0, theta=1e6, eps=0.01))
    bh = ctx.download_bodies()[:, 8:11]
    a64 = orc.direct_f64(b, eps=0.01)
    assert rel_err(bh, a64).max() <= 1e-5


def test_textbook_criterion_converges(ctx, orc):
    b = ics.disk_bulge(20000, seed=5)
    a64 = orc.direct_f64(b, np.arange(0, 20000, 97, dtype=np.uint32), eps=0.01)
    errs = []
    for theta in [1.0, 0.5, 0.25]:
        ctx.upload_bodies(b)
        ctx.accelerate(bh_params(size=128.0, theta=theta, eps=0.01, opening=_lib.OPEN_TEXTBOOK, leaf_size=16))
        a = ctx.download_bodies()[::97, 8:11]
        errs.append(np.median(rel_err(a, a64)))
    print("textbook theta 1, 0.5, 0.25 median errors:", errs)
    assert errs[0] > errs[1] > errs[2] and errs[2] < 1e-3


def _one_octant(n, seed, sign):
    b = ics.plummer(n, seed=seed)
    # every body in one octant of the root cube: the root is then that octant's cell, whose
    # children split on the next digit
    b[:, 0:3] = sign * (np.abs(b[:, 0:3]) * 0.1 + 1.0)
    return b


@pytest.mark.parametrize("case", ["plummer", "root_bucket", "theta0", "octant_first", "octant_last", "outside",
                                  "disk", "cutoff"])
def test_walk_split_matches_single_wave(ctx, case):
    # The walk may split each 64-body group over 2, 4 or 8 waves, each taking a share of
    # the root's octants (the root itself, and lanes that accept it, stay with the first).
    # Every lane's traversal is unchanged, so the applied-node counts are identical and
    # the forces differ only in the order their compensated partial sums are added: a
    # rounding of the parts' sums, which is large next to the total only where the
    # octants' pulls cancel (bodies near the centre of mass).
    size, theta, leaf = 64.0, 0.5, 1
    if case == "plummer":
        b = ics.plummer(20000, seed=31)
    elif case == "root_bucket":
        b, leaf = ics.plummer(12, seed=32), 16
    elif case == "theta0":
        b, theta = ics.plummer(3000, seed=33), 0.0
    elif case.startswith("octant"):
        b = _one_octant(5000, 34, 1.0 if case == "octant_first" else -1.0)
    elif case == "outside":
        b, size = ics.plummer(4000, seed=35), 8.0
    elif case == "disk":
        b, size, leaf = ics.disk_bulge(30000, seed=36), 128.0, 16
    else:
        # the reference's law: no softening, each receiving body's radius cuts off
        b = ics.plummer(8000, seed=38)
        b[:, 3] = np.random.default_rng(38).uniform(0.0, 0.02, len(b)).astype(np.float32)
    eps = 0.0 if case == "cutoff" else 0.01
    p = bh_params(size=size, theta=theta, eps=eps, leaf_size=leaf)
    try:
        got = {}
        for split in (1, 2, 4, 8):
            ctx.bh_configure(split)
            ctx.upload_bodies(b)
            visits, total = ctx.bh_interactions(p)
            ctx.upload_bodies(b)
            ctx.accelerate(p)
            got[split] = (visits.copy(), total, ctx.download_bodies()[:, 8:11].copy())
        v1, t1, a1 = got[1]
        # with a bucket root, or every lane applying the root whole, the later parts have
        # nothing to walk; otherwise they take a share (at 8 parts, one octant each) and the
        # sums' order changes
        idle = case in ("root_bucket", "theta0")
        assert np.array_equal(got[8][2], a1) == idle
        for split in (2, 4, 8):
            v, t, a = got[split]
            assert t == t1 and np.array_equal(v, v1), split
            if idle:
                assert np.array_equal(a, a1), split
            e = rel_err(a, a1)
            print(f"{case} split {split}: rel diff 99% {np.quantile(e, 0.99):.2e} max {e.max():.2e}")
            assert np.quantile(e, 0.99) <= 1e-6 and e.max() <= 1e-5, split
    finally:
        ctx.bh_configure(0)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 65, 129, 1000])
def test_small_and_ragged_systems_every_split(ctx, orc, n):
    # groups that are mostly empty lanes, and a last group cut short, under every split:
    # the traversal counts equal the FP64-centre reference's and the forces match it
    b = ics.plummer(n, seed=40 + n)
    p = bh_params(size=64.0, theta=0.5, eps=0.01)
    x64 = b.copy()
    _, xv = orc.bh_accelerate_f64com(x64, 64.0, theta=0.5, eps=0.01, visits=True)
    try:
        for split in (1, 2, 4, 8):
            ctx.bh_configure(split)
            ctx.upload_bodies(b)
            visits, total = ctx.bh_interactions(p)
            out = ctx.download_bodies()
            assert np.array_equal(visits, xv.astype(visits.dtype)), split
            assert total == visits.sum()
            if n == 1:
                assert np.all(out[:, 8:11] == 0.0)
            else:
                assert rel_err(out[:, 8:11], x64[:, 8:11]).max() <= 1e-5, split
    finally:
        ctx.bh_configure(0)


def test_walk_split_change_recaptures_graph(ctx):
    # a step graph captured under one split is not replayed under another: each replay
    # equals eager stepping under the split in force
    b = ics.disk_bulge(20000, seed=39)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, wrap=1, leaf_size=16)
    try:
        res = {}
        for split in (2, 8):
            ctx.bh_configure(split)
            ctx.upload_bodies(b)
            ctx.step(p, 3)
            eager = ctx.download_bodies()
            ctx.upload_bodies(b)
            ctx.step_graph(p, 3)
            graph = ctx.download_bodies()
            assert np.array_equal(graph.view(np.uint32), eager.view(np.uint32)), split
            res[split] = eager
        assert not np.array_equal(res[2], res[8])
    finally:
        ctx.bh_configure(0)


def test_walk_split_fused_step(ctx):
    # the fused kick-drift runs in the first wave of a group once the parts' sums are in
    b = ics.disk_bulge(6000, seed=37)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, wrap=1, leaf_size=8)
    try:
        out = {}
        for split in (1, 8):
            ctx.bh_configure(split)
            ctx.upload_bodies(b)
            ctx.step(p, 2)
            out[split] = ctx.download_bodies()
        x1, x8 = out[1][:, 0:3].astype(np.float64), out[8][:, 0:3].astype(np.float64)
        assert np.abs(x8 - x1).max() <= 1e-6 * np.abs(x1).max()
        with pytest.raises(_lib.HipError):
            ctx.bh_configure(3)
    finally:
        ctx.bh_configure(0)


def test_step_fused_equals_accelerate_then_integrate(ctx):
    b = ics.disk_bulge(5000, seed=6)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, wrap=1, leaf_size=16)
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    fused = ctx.download_bodies()
    ctx.upload_bodies(b)
    for _ in range(3):
        ctx.accelerate(p)
        ctx.integrate(p)
    split = ctx.download_bodies()
    assert np.array_equal(fused.view(np.uint32), split.view(np.uint32))


@pytest.mark.parametrize("n,leaf,kind,opening", [(20000, 16, "disk", _lib.OPEN_REFERENCE),
                                                 (20000, 5, "disk", _lib.OPEN_REFERENCE),
                                                 (20000, 40, "plummer", _lib.OPEN_REFERENCE),
                                                 (20000, 12, "disk", _lib.OPEN_TEXTBOOK),
                                                 (6000, 16, "clumps", _lib.OPEN_REFERENCE)])
def test_bucket_forces_match_bucket_oracle(ctx, orc, n, leaf, kind, opening):
    # leaf_size > 1: each bucket is tested whole and applied body by body when opened.
    # Against the oracle's bucket walk over the reference tree with FP64 centres: >= 99.99%
    # identical per-body visit counts, and those bodies within 1e-5 (99.9%) / 1e-4 (all).
    # The clumps put hundreds of bodies at spreads down to 1e-6, near the FP32 spacing of
    # their coordinates: the centre of two bodies an odd number of ulps apart is a rounding
    # tie, which the device's FP64 prefix sums round either way, so a few tests there flip
    # (>= 99.9% identical visits); and an aggregate within the softening length moves a
    # force by G M delta / eps^3 with the FP32 rounding delta of its centre, so the clumps
    # run at eps 0.3, where that stays below 5e-6 (measured with FP32 centres in the
    # definition's traversal, tests/bh_definition.py; 9e-4 at eps 0.01).
    size, theta = 64.0, (0.6 if opening == _lib.OPEN_TEXTBOOK else 0.5)
    if kind == "disk":
        b = ics.disk_bulge(n, seed=n + leaf)
        size = 128.0
    elif kind == "plummer":
        b = ics.plummer(n, seed=n + leaf)
    else:
        rng = np.random.default_rng(leaf)
        b = ics.plummer(n, seed=leaf)
        centres = rng.normal(size=(12, 3)).astype(np.float32)
        pick = rng.integers(0, 12, n)
        b[:, 0:3] = centres[pick] + (rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-6, -1, (n, 1))).astype(np.float32)
        b[::37, 0:3] = b[1::37, 0:3][: b[::37].shape[0]]
    eps = 0.3 if kind == "clumps" else 0.01
    ctx.upload_bodies(b)
    p = bh_params(size=size, theta=theta, eps=eps, leaf_size=leaf, opening=opening)
    visits, total = ctx.bh_interactions(p)
    out = ctx.download_bodies()
    x = b.copy()
    _, xv = orc.bh_accelerate_buckets(x, size, theta=theta, leaf_size=leaf, eps=eps,
                                      textbook=opening == _lib.OPEN_TEXTBOOK)
    e = rel_err(out[:, 8:11], x[:, 8:11])
    eq = visits == xv
    print(f"{kind} n={n} leaf={leaf} opening={opening}: same traversal {eq.mean():.5f}, "
          f"99.9% {np.quantile(e[eq], 0.999):.2e} max {e[eq].max():.2e}")
    assert total == visits.sum()
    assert eq.mean() >= (0.999 if kind == "clumps" else 0.9999)
    assert np.quantile(e[eq], 0.999) <= 1e-5 and e[eq].max() <= 1e-4


@pytest.mark.parametrize("leaf,eps", [(1, 0.01), (16, 0.01), (16, 0.0)])
def test_massless_bodies(ctx, orc, leaf, eps):
    # Bodies of zero mass, scattered and filling whole cells.  The library's tree keeps
    # every body and skips every node of zero mass (bhtree.cpp:108-113): leaves and empty
    # nodes carry the threshold -inf, so an empty node always passes and adds exactly zero
    # (its weight is zero; d^2 + eps^2 > 0 wherever the radius cutoff does not remove
    # d^2 = 0), and the interaction count skips it.  Against the tree's definition
    # (tests/bh_definition.py): the same nodes bit for bit, >= 99.99% identical per-body
    # visit counts, those bodies within 1e-5 (99.9%) / 1e-4 (all), every force finite.
    # The reference's own insert treats a leaf of zero mass as empty (bhtree.cpp:26-31):
    # the next body landing there takes its place, while a massless body landing on a
    # populated leaf splits it, so its tree's shape depends on the insertion order.
    # Against the oracle, which restates that tree, the forces keep the same bars on the
    # bodies whose traversal is the same (>= 98% of them here; the others differ only in
    # where cells split, and the massless bodies themselves pull nothing either way).
    n, size, theta = 8000, 64.0, 0.5
    b = ics.plummer(n, seed=leaf + 17)
    b[::5, 7] = 0.0
    b[b[:, 0] > 1.2, 7] = 0.0
    assert 0 < np.count_nonzero(b[:, 7]) < n
    ctx.upload_bodies(b)
    p = bh_params(size=size, theta=theta, eps=eps, leaf_size=leaf)
    nodes = ctx.bh_nodes(p)
    keys, perm = ctx.bh_keys(p)
    spm = b[perm][:, [0, 1, 2, 7]]
    rows = bucket_tree(np.sort(keys), spm, size, theta, leaf)
    thr = np.array([r[3] for r in rows], np.float32)
    assert np.array_equal(nodes["open_sq"].view(np.uint32), thr.view(np.uint32))
    empty = nodes["mass"] == 0
    assert np.any(empty & (nodes["count"] == 0))  # whole internal cells of zero mass
    assert np.all(nodes["open_sq"][empty] == -np.inf)
    visits, total = ctx.bh_interactions(p)
    out = ctx.download_bodies()
    assert total == visits.sum()
    assert np.all(np.isfinite(out[:, 8:11]))
    acc_d, vis_d = walk(rows, spm, eps=eps)
    e = rel_err(out[perm, 8:11], acc_d)
    eq = visits[perm] == vis_d
    print(f"massless leaf={leaf} eps={eps} vs definition: same traversal {eq.mean():.5f}, "
          f"99.9% {np.quantile(e[eq], 0.999):.2e} max {e[eq].max():.2e}")
    assert eq.mean() >= 0.9999
    assert np.quantile(e[eq], 0.999) <= 1e-5 and e[eq].max() <= 1e-4
    x = b.copy()
    _, xv = orc.bh_accelerate_buckets(x, size, theta=theta, leaf_size=leaf, eps=eps)
    e = rel_err(out[:, 8:11], x[:, 8:11])
    eq = visits == xv
    print(f"massless leaf={leaf} eps={eps} vs oracle: same traversal {eq.mean():.5f}, "
          f"99.9% {np.quantile(e[eq], 0.999):.2e} max {e[eq].max():.2e}")
    assert eq.mean() >= 0.98
    assert np.quantile(e[eq], 0.999) <= 1e-5 and e[eq].max() <= 1e-4


def test_bucket_steps_match_bucket_oracle(ctx, orc):
    # k fused Barnes-Hut steps at leaf 16 vs k oracle steps (bucket walk + the reference's
    # integrator) from the same input: positions within 1e-5 for 99.9% of bodies, 1e-4 all
    # (a body whose traversal flips in one step moves by its force change times dt^2)
    n, k, size, dt = 1 << 16, 10, 128.0, 1e-3
    b = ics.disk_bulge(n, seed=99)
    p = bh_params(size=size, theta=0.5, eps=0.01, dt=dt, wrap=0, leaf_size=16)
    ctx.upload_bodies(b)
    ctx.step(p, k)
    gpu = ctx.download_bodies()
    cpu = b.copy()
    for _ in range(k):
        orc.bh_accelerate_buckets(cpu, size, theta=0.5, leaf_size=16, eps=0.01)
        orc.integrate(cpu, np.float32(dt), size, False)
    dx = np.abs(gpu[:, 0:3].astype(np.float64) - cpu[:, 0:3]).max(1)
    print(f"leaf 16 after {k} steps: |dx| 99.9% {np.quantile(dx, 0.999):.2e} max {dx.max():.2e}")
    assert np.quantile(dx, 0.999) <= 1e-5 and dx.max() <= 1e-4


def test_full_size_c3(ctx, orc):
    # N = 2^20 exponential disk + Plummer bulge: keys and permutation bit-exact; forces
    # of leaf size 1 vs the reference Barnes-Hut; forces of leaf size 16 (the configuration
    # the benchmark times) vs the bucket oracle.
    n = 1 << 20
    size = 128.0
    b = ics.disk_bulge(n, seed=2026)
    ctx.upload_bodies(b)
    p = bh_params(size=size, theta=0.5, eps=0.01, leaf_size=16)
    keys, perm = ctx.bh_keys(p)
    ok = orc.morton_keys(b, size, 21)
    assert np.array_equal(keys, ok)
    assert np.array_equal(perm, orc.stable_sort(ok))
    p1 = bh_params(size=size, theta=0.5, eps=0.01, leaf_size=1)
    visits, total = ctx.bh_interactions(p1)
    out = ctx.download_bodies()
    nodes = orc.bh_build(b, size)
    x64 = b.copy()
    _, xv = orc.bh_accelerate_f64com(x64, size, theta=0.5, eps=0.01, visits=True, nodes=nodes)
    ref = b.copy()
    _, rv = orc.bh_accelerate(ref, size, theta=0.5, eps=0.01, visits=True, nodes=nodes)
    e64 = rel_err(out[:, 8:11], x64[:, 8:11])
    eq64 = visits == xv
    e = rel_err(out[:, 8:11], ref[:, 8:11])
    eq = visits == rv
    print(f"C3 leaf 1: vs FP64-centre BH same traversal {eq64.mean():.5f} max {e64[eq64].max():.2e}; "
          f"vs reference same traversal {eq.mean():.5f} max {e[eq].max():.2e} (overall {e.max():.2e}); "
          f"{total / n:.1f} interactions/body")
    assert eq64.mean() >= 0.9999 and e64[eq64].max() <= 1e-4
    assert eq.mean() >= 0.999
    own = rel_err(x64[:, 8:11], ref[:, 8:11])   # the reference's own rounding
    assert np.all(e[eq] <= own[eq] + 1e-4)
    # leaf 16, the benchmarked configuration
    ctx.upload_bodies(b)
    visits, total = ctx.bh_interactions(p)
    out = ctx.download_bodies()
    xb = b.copy()
    _, bv = orc.bh_accelerate_buckets(xb, size, theta=0.5, leaf_size=16, eps=0.01, nodes=nodes)
    eb = rel_err(out[:, 8:11], xb[:, 8:11])
    eqb = visits == bv
    print(f"C3 leaf 16: vs bucket oracle same traversal {eqb.mean():.5f}, 99.9% "
          f"{np.quantile(eb[eqb], 0.999):.2e} max {eb[eqb].max():.2e}; {total / n:.1f} interactions/body")
    assert eqb.mean() >= 0.9999
    assert np.quantile(eb[eqb], 0.999) <= 1e-5 and eb[eqb].max() <= 1e-4
    assert total == visits.sum()


def test_more_than_2p20_bodies(ctx, orc):
    # past 2^20 sorted positions the nearest-smaller searches gain levels above the
    # 4096-position minima: forces of leaf size 1 against the FP64-centre oracle
    n = 1_250_000
    size = 128.0
    b = ics.plummer(n, seed=77)
    ctx.upload_bodies(b)
    p = bh_params(size=size, theta=0.5, eps=0.01, leaf_size=1)
    visits, total = ctx.bh_interactions(p)
    out = ctx.download_bodies()
    x64 = b.copy()
    _, xv = orc.bh_accelerate_f64com(x64, size, theta=0.5, eps=0.01, visits=True)
    e64 = rel_err(out[:, 8:11], x64[:, 8:11])
    eq64 = visits == xv
    print(f"n={n}: same traversal {eq64.mean():.5f}, max {e64[eq64].max():.2e}")
    assert eq64.mean() >= 0.9999 and e64[eq64].max() <= 1e-4
    assert total == visits.sum()


def test_leaf_tuner(ctx):
    b = ics.disk_bulge(50000, seed=12)
    ctx.upload_bodies(b)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, leaf_size=16)
    sizes, ms, best = ctx.bh_tune(p)
    print("leaf sizes", sizes, "ms", np.round(ms, 4), "best", best)
    assert sizes == list(ctx.BH_TUNE_DEFAULT) and np.all(ms > 0)
    assert best == sizes[int(np.argmin(ms))]
    # positions do not move, and the accelerations are the last pass's
    tuned = ctx.download_bodies()
    assert np.array_equal(tuned[:, 0:8].view(np.uint32), b[:, 0:8].view(np.uint32))
    ctx.accelerate(bh_params(size=128.0, theta=0.5, eps=0.01, leaf_size=sizes[-1]))
    assert np.array_equal(ctx.download_bodies().view(np.uint32), tuned.view(np.uint32))
    sizes2, ms2, best2 = ctx.bh_tune(p, [8, 3])
    assert sizes2 == [8, 3] and ms2.size == 2 and best2 in (8, 3)
    with pytest.raises(_lib.HipError):
        ctx.bh_tune(p, [4, 0])


def test_sort_fault_reported_by_every_readback():
    # a radix-sort look-back that gave up waiting leaves a fault word; it is flagged here
    # from the host (no kernel is made to wait) and must surface, once, through each call
    # that brings Barnes-Hut results to the host
    b = ics.disk_bulge(5000, seed=41)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, leaf_size=16)
    with _lib.Context(0) as c:
        c.upload_bodies(b)
        c.step(p, 1)
        for fetch in (lambda: c.download_bodies(), lambda: c.download(), lambda: c.collect_bodies(),
                      lambda: (c.readback(), c.collect_bodies()), lambda: c.synchronize()):
            c.raise_sort_fault()
            with pytest.raises(_lib.HipError) as e:
                fetch()
            assert e.value.status == _lib.ERR_HIP and "radix-sort" in str(e.value)
            fetch()  # reported once: the next call succeeds


def test_nlogn_refuses_more_than_2p30_minus_1_bodies():
    # the sort's look-back words hold 30-bit totals; a larger build is refused before any
    # kernel runs (the 2^30-body arrays are allocated, never touched)
    with _lib.Context(0) as c:
        c.reserve(_lib.MAX_NLOGN_BODIES + 1)
        for call in (lambda: c.accelerate(bh_params(size=128.0)), lambda: c.bh_keys(bh_params(size=128.0)),
                     lambda: c.step_graph(bh_params(size=128.0, dt=1e-3), 2)):
            with pytest.raises(_lib.HipError) as e:
                call()
            assert e.value.status == _lib.ERR_UNSUPPORTED
        with pytest.raises(_lib.HipError) as e:
            c.reserve(_lib.MAX_BODIES + 1)
        assert e.value.status == _lib.ERR_INVALID


def test_profile_steps_equal_plain_steps(ctx):
    # nbhip_bh_profile runs the same fused steps as nbhip_step, with events around the phases
    b = ics.disk_bulge(30000, seed=43)
    p = bh_params(size=128.0, theta=0.5, eps=0.01, dt=1e-3, wrap=1, leaf_size=16)
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    plain = ctx.download_bodies()
    ctx.upload_bodies(b)
    build_ms, walk_ms = ctx.bh_profile(p, 3)
    prof = ctx.download_bodies()
    assert np.array_equal(prof.view(np.uint32), plain.view(np.uint32))
    assert build_ms > 0 and walk_ms > 0
    with pytest.raises(_lib.HipError):
        ctx.bh_profile(_lib.params(size=128.0), 1)   # direct-sum params


def auto_root(b):
    # the root cube found from the bodies: smallest power of two >= 2 max |coordinate|
    m = float(np.abs(b[:, 0:3].astype(np.float64)).max())
    return 1.0 if m == 0.0 else float(2.0 ** np.ceil(np.log2(2.0 * m)))


@pytest.mark.parametrize("scale", [1.0, 37.5, 1e-3])
def test_root_found_from_the_bodies(ctx, orc, scale):
    # size 0 in NLogN mode: the root edge comes from an FP64 reduction on the device; keys,
    # permutation and forces equal those of an explicit root of that edge, bit for bit,
    # and the forces match the bucket oracle there
    b = ics.disk_bulge(30000, seed=int(scale * 10) + 3)
    b[:, 0:3] *= np.float32(scale)
    b[5, 0:3] = [0.0, -np.float32(24.0 * scale), 0.0]   # a far body sets the extent along -y
    edge = auto_root(b)
    ctx.upload_bodies(b)
    p0 = bh_params(size=0.0, theta=0.5, eps=0.01 * scale, leaf_size=16)
    pe = bh_params(size=edge, theta=0.5, eps=0.01 * scale, leaf_size=16)
    k0, perm0 = ctx.bh_keys(p0)
    ke, perme = ctx.bh_keys(pe)
    assert np.array_equal(k0, ke) and np.array_equal(perm0, perme)
    assert np.array_equal(k0, orc.morton_keys(b, edge, 21))
    ctx.accelerate(p0)
    assert ctx.bh_root() == edge
    a0 = ctx.download_bodies()
    tree0 = ctx.bh_tree()
    ctx.accelerate(pe)
    ae = ctx.download_bodies()
    assert np.array_equal(a0.view(np.uint32), ae.view(np.uint32))
    assert tree0["size"][0] == edge and np.array_equal(tree0, ctx.bh_tree())
    x = b.copy()
    _, xv = orc.bh_accelerate_buckets(x, edge, theta=0.5, leaf_size=16, eps=0.01 * scale)
    visits, _ = ctx.bh_interactions(p0)
    eq = visits == xv
    e = rel_err(a0[:, 8:11], x[:, 8:11])
    assert eq.mean() >= 0.9999 and np.quantile(e[eq], 0.999) <= 1e-5 and e[eq].max() <= 1e-4
    # fused steps with size 0 find the root again each step and never wrap: the same as an
    # explicit root of that edge without wrap, while the extent stays within it
    dt = 1e-3 * scale ** 1.5   # the same fraction of a dynamical time at every scale
    ctx.upload_bodies(b)
    ctx.step(bh_params(size=0.0, theta=0.5, eps=0.01 * scale, dt=dt, wrap=1, leaf_size=16), 2)
    s0 = ctx.download_bodies()
    ctx.upload_bodies(b)
    ctx.step(bh_params(size=edge, theta=0.5, eps=0.01 * scale, dt=dt, wrap=0, leaf_size=16), 2)
    assert np.array_equal(s0.view(np.uint32), ctx.download_bodies().view(np.uint32))
    with pytest.raises(_lib.HipError):
        ctx.accelerate(bh_params(size=-1.0))


def test_key_order_between_steps_is_invisible(ctx):
    # Steps leave the bodies in key order on the device.  A dense cluster under the
    # reference's root edge (10000: 21-level cells 0.005 wide) puts many bodies in shared
    # finest cells, opened and applied body by body, and bodies change cells between steps,
    # so the last build's key order often disagrees with the caller's order among equal
    # keys; every build puts those in the caller's order.  Stepping on therefore equals
    # re-uploading in the caller's order before every step, bit for bit, eagerly and from
    # a graph -- and the order among equal keys does matter: the same bodies listed in
    # reverse give different bits.
    n = 20000
    b = ics.plummer(n, seed=3, a=0.02)
    p = bh_params(size=10000.0, theta=0.5, dt=2e-3, eps=1e-3, leaf_size=1, wrap=0)
    steps = 4

    def fresh_each_step(bodies):
        x = bodies
        for _ in range(steps):
            ctx.upload_bodies(x)
            ctx.step(p, 1)
            x = ctx.download_bodies()
        return x

    ctx.upload_bodies(b)
    keys, _ = ctx.bh_keys(p)
    assert np.unique(keys).size < n - 2000   # many bodies share a finest cell
    ctx.step(p, steps)
    on = ctx.download_bodies()
    ctx.upload_bodies(b)
    ctx.step_graph(p, steps)
    graph = ctx.download_bodies()
    ref = fresh_each_step(b)
    assert np.array_equal(on, ref)
    assert np.array_equal(graph, ref)
    rev = fresh_each_step(b[::-1])[::-1]
    assert not np.array_equal(rev[:, 8:11], ref[:, 8:11])


def test_caller_order_calls_after_steps(ctx):
    # partial downloads, partial uploads, staged readback and diagnostics address bodies by
    # the caller's index whatever order the steps left on the device
    n = 30000
    b = ics.disk_bulge(n, seed=8)
    p = bh_params(size=128.0, eps=0.01, leaf_size=16, wrap=0)
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    full = ctx.download_bodies()
    ctx.upload_bodies(full)
    d_ref = ctx.diagnostics(p)
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    assert np.array_equal(ctx.diagnostics(p), d_ref)
    ctx.step(p, 0)
    pm, vr, acc = ctx.download(offset=1000, n=500)
    assert np.array_equal(pm[:, 0:3], full[1000:1500, 0:3]) and np.array_equal(pm[:, 3], full[1000:1500, 7])
    assert np.array_equal(acc[:, 0:3], full[1000:1500, 8:11])
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    ctx.readback()
    staged, was_staged = ctx.collect_bodies()
    assert was_staged and np.array_equal(staged, full)
    # a partial upload after steps changes exactly the bodies it names
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    new_pm = np.zeros((10, 4), np.float32)
    new_pm[:, 0] = np.arange(10)
    new_pm[:, 3] = 1.0
    ctx.upload(pos_mass=new_pm, offset=20, n=10)
    after = ctx.download_bodies()
    want = full.copy()
    want[20:30, 0:3] = new_pm[:, 0:3]
    want[20:30, 7] = 1.0
    assert np.array_equal(after, want)
