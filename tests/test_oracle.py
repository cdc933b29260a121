"""Pins the CPU oracle (oracle/nbody_oracle.c) before anything is checked against it.

Each oracle function is checked against the golden vectors written from the reference's
own sources (tests/golden/make_golden.py), and, where those leave something open, against
what the mathematics fixes: closed forms, invariants, special cases and brute force.
"""
import os

import numpy as np
import pytest

from nbody_amd import ics

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# ---- golden vectors (bit-exact) --------------------------------------------------------

def test_gravity_pairs_match_reference(orc):
    g = gold("gravity_pairs.npz")
    for i in range(g["pos"].shape[0]):
        out = orc.gravity_pair(g["pos"][i], g["radius"][i], g["src"][i], g["mass"][i], g["G"][i], 0.0)
        assert np.array_equal(bits(out), bits(g["acc"][i])), i


def test_wrap_matches_reference(orc):
    g = gold("wrap_cases.npz")
    for x, s, w in zip(g["x"], g["size"], g["wrapped"]):
        assert np.float32(orc.wrap(float(x), float(s))).view(np.uint32) == np.float32(w).view(np.uint32)


@pytest.mark.parametrize("case,G", [("disk", 1.0), ("cube", 4.0)])
def test_brute_force_matches_reference(orc, case, G):
    g = gold("brute_force.npz")
    b = g[case].copy()
    orc.brute_force(b, G=G, eps=0.0, threads=3)
    assert np.array_equal(bits(b), bits(g[case + "_out"]))


def test_integrate_matches_reference(orc):
    g = gold("integrate.npz")
    for out, dt, size, wrap in [("out_a", 1 / 60, 10000.0, True), ("out_b", 0.5, 50.0, True),
                                ("out_c", 0.5, 50.0, False)]:
        b = g["bodies"].copy()
        orc.integrate(b, np.float32(dt), size, wrap)
        assert np.array_equal(bits(b), bits(g[out])), out


@pytest.mark.parametrize("case,size", [("disk64", 10000.0), ("cube100", 1024.0)])
def test_tree_matches_reference(orc, case, size):
    g = gold("tree.npz")
    nodes = orc.bh_build(g[case], size)
    assert np.array_equal(nodes.view(np.uint8), g[case + "_nodes"])


@pytest.mark.parametrize("theta", [0.5, 8.0])
def test_barnes_hut_matches_reference(orc, theta):
    g = gold("barnes_hut.npz")
    tag = "05" if theta == 0.5 else "8"
    b = g["bodies"].copy()
    _, visits = orc.bh_accelerate(b, 10000.0, theta=theta, visits=True, threads=2)
    assert np.array_equal(bits(b), bits(g["out_theta" + tag]))
    assert np.array_equal(visits, g["visits_theta" + tag])


@pytest.mark.parametrize("size", [256, 10000, 100])
def test_morton_keys_match_reference_descent(orc, size):
    g = gold("cell_keys.npz")
    keys = orc.morton_keys(g["bodies"], float(size), 21)
    assert np.array_equal(keys, g["keys_%d" % size])


# ---- mathematics (what the golden vectors leave open) ---------------------------------

def test_two_body_closed_form(orc):
    # |a| = G m / r^2 along the separation; softened: G m r / (r^2 + eps^2)^{3/2}
    a = orc.gravity_pair([0, 0, 0], 0.0, [3, 0, 4], 2.0, G=1.5, eps=0.0)
    assert np.allclose(a, 1.5 * 2.0 / 25.0 * np.array([0.6, 0, 0.8]), rtol=1e-6)
    a = orc.gravity_pair([0, 0, 0], 0.0, [3, 0, 4], 2.0, G=1.0, eps=0.5)
    assert np.allclose(a, 2.0 * np.array([3, 0, 4]) / (25.25 ** 1.5), rtol=1e-6)
    # cutoff on the receiving body's radius (not strict: equal distance is cut)
    assert np.all(orc.gravity_pair([0, 0, 0], 5.0, [3, 0, 4], 2.0) == 0)
    assert np.any(orc.gravity_pair([0, 0, 0], 4.99, [3, 0, 4], 2.0) != 0)


def test_brute_force_momentum_and_f64(orc):
    from nbody_amd import ics
    b = ics.plummer(512, seed=7)
    orc.brute_force(b, eps=0.01)
    m = b[:, 7:8].astype(np.float64)
    a = b[:, 8:11].astype(np.float64)
    # Newton's third law: sum m a = 0 up to FP32 rounding
    assert np.abs((m * a).sum(0)).max() < 1e-5 * np.abs(m * a).sum(0).max()
    a64 = orc.direct_f64(b, eps=0.01)
    rel = np.linalg.norm(a - a64, axis=1) / np.linalg.norm(a64, axis=1)
    assert rel.max() < 1e-4 and np.median(rel) < 1e-6
    # FP64 sample path equals the full path
    s = np.array([0, 17, 511], np.uint32)
    assert np.array_equal(orc.direct_f64(b, s, eps=0.01), a64[s])


def test_direct_f64_plummer_field(orc):
    # Outside a uniform-density ball the field is G M / r^2 (shell theorem): a body far
    # away sees the whole cloud as a point mass at its centre.
    rng = np.random.default_rng(3)
    n = 4000
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = rng.uniform(0, 1, n) ** (1 / 3)
    b = np.zeros((n + 1, 12), np.float32)
    b[:n, 0:3] = u * r[:, None]
    b[:n, 7] = 1.0 / n
    b[n, 0:3] = [50.0, 0, 0]
    b[n, 7] = 1e-9
    a = orc.direct_f64(b, np.array([n], np.uint32))[0]
    assert abs(a[0] + 1.0 / 2500.0) < 1e-3 / 2500.0


def test_integrate_closed_form(orc):
    # G = 0 drift: x + v dt exactly; wrap 499 + 10 -> -491 in a box of 1000
    # (tests/source/test_wrap.cpp:15-26,48)
    b = np.zeros((1, 12), np.float32)
    b[0, 0:3] = [499, 0, 0]
    b[0, 4:7] = [10, 0, 0]
    orc.integrate(b, 1.0, 1000.0, True)
    assert b[0, 0] == -491.0
    b[0, 0:3] = [499, 0, 0]
    orc.integrate(b, 1.0, 1000.0, False)
    assert b[0, 0] == 509.0


def test_energy_f64_two_bodies(orc):
    b = np.zeros((2, 12), np.float32)
    b[0, 0:3] = [1, 0, 0]
    b[1, 0:3] = [-1, 0, 0]
    b[0, 4:7] = [0, 0.5, 0]
    b[1, 4:7] = [0, -0.5, 0]
    b[:, 7] = [1.0, 1.0]
    e = orc.energy_f64(b, G=1.0, eps=0.0)
    assert np.isclose(e[0], 0.25)          # 2 * 1/2 * 1 * 0.25
    assert np.isclose(e[1], -0.5)          # -G m m / r
    assert np.allclose(e[2:5], 0)
    assert np.isclose(e[7], 1.0)           # L_z = 2 * (1 * 0.5)
    assert e[8] == 2.0


def test_stable_sort_is_std_stable_sort(orc):
    rng = np.random.default_rng(11)
    keys = rng.integers(0, 50, 5000).astype(np.uint64)  # many ties
    perm = orc.stable_sort(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))


def test_morton_descent_is_floor_quantisation(orc):
    # With a power-of-two box the comparison descent equals floor quantisation, digit =
    # complement of the usual Morton bit (a set bit means the lower half).
    rng = np.random.default_rng(5)
    b = np.zeros((3000, 12), np.float32)
    b[:, 0:3] = rng.uniform(-64, 64, (3000, 3)).astype(np.float32)
    keys = orc.morton_keys(b, 128.0, 10)
    cell = np.floor((b[:, 0:3].astype(np.float64) + 64.0) / (128.0 / 1024)).astype(np.int64)
    expect = np.zeros(3000, np.uint64)
    for lvl in range(10):
        bit = 9 - lvl
        q = ((1 - ((cell[:, 0] >> bit) & 1)) | ((1 - ((cell[:, 1] >> bit) & 1)) << 1)
             | ((1 - ((cell[:, 2] >> bit) & 1)) << 2))
        expect = (expect << np.uint64(3)) | q.astype(np.uint64)
    assert np.array_equal(keys, expect)


def test_bh_special_cases(orc):
    from nbody_amd import ics
    b = ics.plummer(300, seed=2)
    # theta -> 0 (reference convention): only leaves can be applied... and the root is
    # accepted only if dist^2 > 0; theta huge: every internal node is opened, so the sum
    # visits every body's leaf = brute force up to summation order.
    bf = b.copy()
    orc.brute_force(bf, eps=0.0)
    bh = b.copy()
    orc.bh_accelerate(bh, 64.0, theta=1e6, eps=0.0)
    rel = np.linalg.norm(bh[:, 8:11] - bf[:, 8:11], axis=1) / np.linalg.norm(bf[:, 8:11], axis=1)
    assert rel.max() < 1e-4
    # theta = 0 accepts any internal node with a distance > 0: for a body outside the
    # root's com every accepted node is the root itself -> a = G M (com - x) / |.|^3
    nodes = orc.bh_build(b, 64.0)
    one = b.copy()
    orc.bh_accelerate(one, 64.0, theta=0.0, nodes=nodes)
    com = nodes[0]["com"].astype(np.float64)
    d = com - one[:, 0:3].astype(np.float64)
    expect = nodes[0]["mass"] * d / np.linalg.norm(d, axis=1, keepdims=True) ** 3
    assert np.allclose(one[:, 8:11], expect, rtol=1e-4)
    # mass conservation in the tree (tests/source/test_bhtree.cpp:169-212)
    assert np.isclose(nodes[0]["mass"], b[:, 7].sum(), rtol=1e-5)


def test_bh_f64com_oracle(orc):
    # Pins the FP64-centre variant: same traversal as the reference on the golden disk,
    # forces within the reference's own FP32 rounding, and with every cell opened it is
    # the FP64 direct sum.
    g = np.load(os.path.join(GOLD, "barnes_hut.npz"))
    b = g["bodies"].copy()
    _, v = orc.bh_accelerate_f64com(b, 10000.0, theta=0.5, visits=True)
    assert np.array_equal(v, g["visits_theta05"])
    ref = g["out_theta05"][:, 8:11].astype(np.float64)
    rel = np.linalg.norm(b[:, 8:11] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert rel.max() < 1e-4
    from nbody_amd import ics
    p = ics.plummer(1500, seed=4)
    x = p.copy()
    orc.bh_accelerate_f64com(x, 64.0, theta=1e6, eps=0.01)
    a64 = orc.direct_f64(p, eps=0.01)
    assert (np.linalg.norm(x[:, 8:11] - a64, axis=1) / np.linalg.norm(a64, axis=1)).max() < 1e-6


def test_o3_build_is_bit_identical(orc):
    # the CPU baseline's second column (-O3 -march=x86-64-v3, contraction still off) must
    # compute exactly what the -O2 oracle does
    b = ics.disk_bulge(3000, seed=4)
    runs = []
    for v in ("o2", "o3"):
        orc.variant(v)
        try:
            x = b.copy()
            orc.brute_force(x, eps=0.01)
            y = b.copy()
            _, vis = orc.bh_accelerate(y, 128.0, theta=0.5, eps=0.01, visits=True)
            runs.append((x, y, vis, orc.morton_keys(b, 128.0, 21)))
        finally:
            orc.variant("o2")
    for a, c in zip(runs[0], runs[1]):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(c).view(np.uint8))


# ---- the bucket-aware Barnes-Hut oracle (leaf_size > 1) ---------------------------------

def _clumps(n, seed):
    # many bodies in a few deep cells (long one-child chains) and exact duplicates (cells
    # still holding several bodies at 21 levels)
    rng = np.random.default_rng(seed)
    b = ics.plummer(n, seed=seed)
    centres = rng.normal(size=(12, 3)).astype(np.float32)
    pick = rng.integers(0, 12, n)
    b[:, 0:3] = centres[pick] + (rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-6, -1, (n, 1))).astype(np.float32)
    b[::37, 0:3] = b[1::37, 0:3][: b[::37].shape[0]]
    return b


def test_bucket_oracle_leaf1_is_the_f64com_oracle(orc):
    # one body per leaf and no two bodies in a 21-level cell: the reference traversal with
    # FP64 centres, visit for visit (the two oracles sum the centres in different orders)
    for b, size in [(ics.disk_bulge(20000, seed=3), 128.0), (ics.plummer(5000, seed=9), 64.0)]:
        keys = orc.morton_keys(b, size, 21)
        assert np.unique(keys).size == keys.size
        for theta in (0.5, 2.0):
            x1 = b.copy()
            _, v1 = orc.bh_accelerate_f64com(x1, size, theta=theta, eps=0.01, visits=True)
            x2 = b.copy()
            _, v2 = orc.bh_accelerate_buckets(x2, size, theta=theta, leaf_size=1, eps=0.01)
            assert np.array_equal(v1, v2)
            ref = x1[:, 8:11].astype(np.float64)
            rel = np.linalg.norm(x2[:, 8:11] - ref, axis=1) / np.linalg.norm(ref, axis=1)
            assert rel.max() <= 1e-6


@pytest.mark.parametrize("leaf", [1, 4, 16, 40])
def test_bucket_oracle_every_cell_open_is_the_direct_sum(orc, leaf):
    # theta -> infinity opens every cell and every bucket: each body visits every body once
    b = ics.disk_bulge(6000, seed=leaf)
    x = b.copy()
    _, v = orc.bh_accelerate_buckets(x, 128.0, theta=1e6, leaf_size=leaf, eps=0.01)
    assert np.all(v == b.shape[0])
    a64 = orc.direct_f64(b, eps=0.01)
    assert (np.linalg.norm(x[:, 8:11] - a64, axis=1) / np.linalg.norm(a64, axis=1)).max() <= 2e-7


@pytest.mark.parametrize("leaf", [1, 2, 5, 16, 17, 40])
@pytest.mark.parametrize("kind,textbook", [("disk", False), ("clumps", False), ("disk", True)])
def test_bucket_oracle_matches_definition(orc, leaf, kind, textbook):
    # the oracle (the reference's insertion tree, marked with buckets) against a traversal
    # of the bucket tree written from its definition over key prefixes (tests/bh_definition.py):
    # same visits body for body, forces equal to FP64 rounding
    from bh_definition import bucket_tree, walk
    n, size = 3000, 64.0
    b = ics.disk_bulge(n, seed=20 + leaf) if kind == "disk" else _clumps(n, 20 + leaf)
    theta = 0.6 if textbook else 0.7
    x = b.copy()
    _, v = orc.bh_accelerate_buckets(x, size, theta=theta, leaf_size=leaf, eps=0.01, textbook=textbook)
    keys = orc.morton_keys(b, size, 21)
    perm = orc.stable_sort(keys)
    spm = b[perm][:, [0, 1, 2, 7]]
    rows = bucket_tree(keys[perm], spm, size, theta, leaf, textbook=textbook)
    acc, dv = walk(rows, spm, eps=0.01)
    assert np.array_equal(v[perm], dv)
    a = x[perm, 8:11].astype(np.float64)
    rel = np.linalg.norm(a - acc, axis=1) / np.linalg.norm(acc, axis=1)
    assert rel.max() <= 2e-7   # the oracle's FP32 output
    if leaf > 2 and kind == "disk":
        # buckets of three or more bodies change what is applied (an opened pair is its
        # two leaves either way): the comparison would fail for a leaf-1 traversal
        _, v1 = orc.bh_accelerate_buckets(b.copy(), size, theta=theta, leaf_size=1, eps=0.01, textbook=textbook)
        assert not np.array_equal(v1, v)


def _close_groups(n, size, seed, k=6):
    # groups of k bodies 2^-21 .. 2^-18 root edges from the first of the group: many groups
    # share a cell 18 levels deep and split further down by 21 levels
    rng = np.random.default_rng(seed)
    b = ics.plummer(n, seed=seed)
    lead = (np.arange(n) // k) * k
    d = rng.normal(size=(n, 3)) * (size * 2.0 ** rng.uniform(-21, -18, (n, 1)))
    d[lead == np.arange(n)] = 0.0
    b[:, 0:3] = (b[lead, 0:3] + d).astype(np.float32)
    return b


@pytest.mark.parametrize("leaf", [1, 2, 16])
def test_bucket_oracle_18_levels_matches_definition(orc, leaf):
    # the oracle's 18-level buckets (a batch of systems) against the definition's tree cut
    # at 18 levels: same visits body for body, forces equal to FP64 rounding; and the input
    # does put several bodies in 18-level cells that 21 levels would split, which changes
    # the visits
    from bh_definition import bucket_tree, walk
    n, size, theta = 3000, 64.0, 0.7
    b = _close_groups(n, size, 40 + leaf, k=6 if leaf < 4 else 3 * leaf)   # groups above the bucket size
    keys = orc.morton_keys(b, size, 21)
    k18 = keys >> np.uint64(9)
    assert np.unique(k18).size < np.unique(keys).size - 100
    x = b.copy()
    _, v = orc.bh_accelerate_buckets(x, size, theta=theta, leaf_size=leaf, eps=0.01, max_level=18)
    perm = orc.stable_sort(keys)
    spm = b[perm][:, [0, 1, 2, 7]]
    rows = bucket_tree(keys[perm], spm, size, theta, leaf, levels=18)
    acc, dv = walk(rows, spm, eps=0.01)
    assert np.array_equal(v[perm], dv)
    a = x[perm, 8:11].astype(np.float64)
    rel = np.linalg.norm(a - acc, axis=1) / np.linalg.norm(acc, axis=1)
    assert rel.max() <= 2e-7
    _, v21 = orc.bh_accelerate_buckets(b.copy(), size, theta=theta, leaf_size=leaf, eps=0.01)
    assert not np.array_equal(v21, v)


def test_bucket_oracle_refuses_bad_levels(orc):
    b = ics.plummer(100, seed=1)
    for lv in (0, 22):
        with pytest.raises(MemoryError):
            orc.bh_accelerate_buckets(b.copy(), 64.0, leaf_size=4, max_level=lv)


def test_bucket_oracle_catches_a_shallow_threshold(orc):
    # a planted mistake in the definition's walk -- a bucket tested at the edge of the
    # shallowest cell of its chain instead of the deepest -- must change the visits on
    # clumpy input, so the comparison above would catch the same slip in the oracle
    from bh_definition import bucket_tree, walk
    n, size, theta, leaf = 3000, 64.0, 0.7, 16
    b = _clumps(n, 7)
    keys = orc.morton_keys(b, size, 21)
    perm = orc.stable_sort(keys)
    spm = b[perm][:, [0, 1, 2, 7]]
    rows = bucket_tree(keys[perm], spm, size, theta, leaf)
    _, good = walk(rows, spm, eps=0.01)
    bad_rows = [list(r) for r in rows]
    for r in bad_rows:
        if r[1] > 1:
            r[3] = np.float32(r[3] * np.float32(4.0))   # one level shallower
    _, bad = walk(bad_rows, spm, eps=0.01)
    assert not np.array_equal(good, bad)
    _, v = orc.bh_accelerate_buckets(b.copy(), size, theta=theta, leaf_size=leaf, eps=0.01)
    assert np.array_equal(v[perm], good)
