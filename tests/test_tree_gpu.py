"""The device octree exported in the reference's bh::Node layout (nbhip_bh_tree), through
the C ABI, against the oracle's restatement of the reference's tree build
(source/detail/tree.h:14-29, source/bhtree.cpp:14-95).

Bars (stated per check):
  * structure: the same cells in the same depth-first child order (children + q), the
    same node count, bit-exact bounds (centre and edge), the same leaf/internal split,
    single-body leaves bit-exact (position and mass), and `next` threaded as the
    reference threads it (child q < 7 -> its sibling, child 7 -> the parent's next);
  * internal masses and centres of mass: within one FP32 rounding (relative 2^-23 of
    the magnitude) of the FP64 sums over the cell's bodies, computed here from the
    oracle tree's own leaves.  The reference's incremental FP32 means differ from those
    sums by their own rounding (up to ~1e-5 of the cell edge at 2000 bodies), so it is
    compared with that looser bound;
  * the reference traversal (Tree::apply, source/bhtree.cpp:97-138, as the oracle
    restates it) over the exported tree: with centres recomputed in FP64 from the leaves,
    bit-identical to the same over the oracle's tree (same tree, same order); with the
    exported FP32 centres, >= 99.9% of bodies open the same nodes as the FP64-centre
    oracle and those forces agree with it to 1e-5 (99.9%) / 1e-4 (all) (a CPU emulation
    of the export measured 2.7e-6 / 1.4e-5 on this input, the reference's own tree
    6.3e-5 / 1.4e-4);
  * known layouts from the reference's own tree tests (tests/source/test_bhtree.cpp:30-214).
"""
import numpy as np
import pytest

from nbody_amd import _lib, ics

pytestmark = pytest.mark.gpu


def bodies_of(pos, mass):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    b = np.zeros((pos.shape[0], 12), np.float32)
    b[:, 0:3] = pos
    b[:, 7] = mass
    return b


def tree_params(size, **kw):
    kw.setdefault("mode", _lib.MODE_NLOGN)
    return _lib.params(size=size, **kw)


def canonical(nodes):
    """Node indices in depth-first order following children + q (the reference's own
    traversal order), with each node's depth."""
    order, depth = [], []
    stack = [(0, 0)]
    while stack:
        i, d = stack.pop()
        order.append(i)
        depth.append(d)
        c = int(nodes[i]["children"])
        if c:
            for q in range(7, -1, -1):
                stack.append((c + q, d + 1))
    return np.array(order), np.array(depth)


def expected_next(nodes):
    """`next` of every node by the reference's rule (source/bhtree.cpp:44-54)."""
    nxt = np.zeros(len(nodes), np.int64)
    stack = [(0, 0)]
    while stack:
        i, n = stack.pop()
        nxt[i] = n
        c = int(nodes[i]["children"])
        if c:
            for q in range(8):
                stack.append((c + q, c + q + 1 if q < 7 else n))
    return nxt


def f64_sums(nodes, order):
    """FP64 mass and centre of mass of every node from the single-body leaves below it."""
    m = np.zeros(len(nodes))
    mx = np.zeros((len(nodes), 3))
    for i in order[::-1]:  # children before parents
        c = int(nodes[i]["children"])
        if c:
            m[i] = m[c:c + 8].sum()
            mx[i] = mx[c:c + 8].sum(axis=0)
        else:
            m[i] = float(nodes[i]["mass"])
            mx[i] = m[i] * nodes[i]["com"].astype(np.float64)
    com = np.where(m[:, None] > 0, mx / np.where(m > 0, m, 1)[:, None], 0.0)
    return m, com


def compare_structure(gpu, ref):
    assert len(gpu) == len(ref)
    og, dg = canonical(gpu)
    orf, dr = canonical(ref)
    assert len(og) == len(gpu) and len(orf) == len(ref)  # every node reachable once
    assert np.array_equal(dg, dr)
    g, r = gpu[og], ref[orf]
    assert np.array_equal(g["center"].view(np.uint32), r["center"].view(np.uint32))
    assert np.array_equal(g["size"].view(np.uint32), r["size"].view(np.uint32))
    assert np.array_equal(g["children"] != 0, r["children"] != 0)
    assert np.array_equal(gpu["next"], expected_next(gpu))
    assert np.array_equal(ref["next"], expected_next(ref))
    assert not gpu["pad0"].any() and not gpu["pad1"].any()
    leaf = (r["children"] == 0) & (r["mass"] != 0)
    assert np.array_equal(g["mass"][leaf].view(np.uint32), r["mass"][leaf].view(np.uint32))
    assert np.array_equal(g["com"][leaf].view(np.uint32), r["com"][leaf].view(np.uint32))
    empty = r["mass"] == 0
    assert np.all(g["mass"][empty] == 0) and np.all(g["com"][empty] == 0)
    return og, orf


def test_known_layouts(ctx):
    # tests/source/test_bhtree.cpp:30-142 -- one body: the root is the leaf
    ctx.upload_bodies(bodies_of([50, 50, 50], 1.0))
    t = ctx.bh_tree(tree_params(100.0))
    assert len(t) == 1 and t[0]["mass"] == 1 and tuple(t[0]["com"]) == (50, 50, 50)
    assert t[0]["children"] == 0 and t[0]["next"] == 0 and t[0]["size"] == 100
    # two, then three bodies in different octants: 9 nodes, child q at 1 + q
    for k, mass in ((2, 2.0), (3, 3.0)):
        pos = [[50, 50, 50], [-50, 50, 50], [50, -50, 50]][:k]
        ctx.upload_bodies(bodies_of(pos, 1.0))
        t = ctx.bh_tree(tree_params(100.0))
        assert len(t) == 9 and t[0]["mass"] == mass and t[0]["children"] == 1
        for q, p in enumerate(pos):
            assert t[1 + q]["mass"] == 1 and tuple(t[1 + q]["com"]) == tuple(p)
            assert t[1 + q]["children"] == 0 and t[1 + q]["size"] == 50
        if k == 2:
            assert tuple(t[0]["com"]) == (0, 50, 50)
        assert list(t["next"][1:9]) == [2, 3, 4, 5, 6, 7, 8, 0]
    # two bodies in one octant of the root: it splits once more (1 + 8 + 8 nodes)
    ctx.upload_bodies(bodies_of([[10, 10, 10], [40, 40, 40]], [1.0, 3.0]))
    t = ctx.bh_tree(tree_params(100.0))
    assert len(t) == 17 and t[0]["mass"] == 4
    c = t[1]["children"]
    assert c == 9 and t[1]["mass"] == 4
    q0 = (10 < 25) | (10 < 25) << 1 | (10 < 25) << 2
    q1 = 0
    assert t[c + q0]["mass"] == 1 and tuple(t[c + q0]["com"]) == (10, 10, 10)
    assert t[c + q1]["mass"] == 3 and tuple(t[c + q1]["com"]) == (40, 40, 40)
    # coincident and near-coincident pairs terminate, mass 2 at the root
    # (tests/source/test_bhtree.cpp:144-165)
    for second in ([1, 1, 1], [1 + np.finfo(np.float32).eps, 1, 1]):
        ctx.upload_bodies(bodies_of([[1, 1, 1], second], 1.0))
        t = ctx.bh_tree(tree_params(100.0))
        assert len(t) > 1 and t[0]["mass"] == 2
        assert np.allclose(t[0]["com"], [1, 1, 1], rtol=1e-5)


def test_mass_conserved_100000(ctx):
    # tests/source/test_bhtree.cpp:189-212: normal(0, 0.5) clamped to [-1, 1] x 100, root 200
    rng = np.random.default_rng(7)
    pos = np.clip(rng.normal(0, 0.5, (100000, 3)), -1, 1) * 100
    ctx.upload_bodies(bodies_of(pos, 1.0))
    t = ctx.bh_tree(tree_params(200.0))
    assert t[0]["mass"] == np.float32(100000)
    assert len(t) > 1
    leaves = t[(t["children"] == 0) & (t["mass"] > 0)]
    assert leaves["mass"].sum(dtype=np.float64) == 100000.0


@pytest.mark.parametrize("kind,n,size", [("plummer", 300, 64.0), ("disk", 2000, 128.0),
                                         ("cube", 512, 1024.0), ("plummer", 20000, 64.0),
                                         ("disk", 50000, 128.0)])
def test_structure_vs_oracle(ctx, orc, kind, n, size):
    b = {"plummer": lambda: ics.plummer(n, seed=n),
         "disk": lambda: ics.disk_bulge(n, seed=n),
         "cube": lambda: ics.cube(n, seed=n)}[kind]()
    ctx.upload_bodies(b)
    gpu = ctx.bh_tree(tree_params(size))
    ref = orc.bh_build(b, size)
    og, orf = compare_structure(gpu, ref)
    # internal cells: one FP32 rounding from the FP64 sums over the cell's bodies
    m64, c64 = f64_sums(ref, orf)
    g = gpu[og]
    inner = ref[orf]["children"] != 0
    mg = g["mass"][inner].astype(np.float64)
    mt = m64[orf][inner]
    assert np.all(np.abs(mg - mt) <= 2.0 ** -23 * mt)
    cg = g["com"][inner].astype(np.float64)
    ct = c64[orf][inner]
    # (plus 1e-10 absolute: the FP64 prefix sums lose ~1e-11 to cancellation near 0)
    assert np.all(np.abs(cg - ct) <= 2.0 ** -23 * np.abs(ct) + 1e-10)
    # the reference's incremental FP32 means: within its own rounding
    r = ref[orf]
    rel_m = np.abs(r["mass"][inner] - mt) / mt
    dc = np.abs(r["com"][inner].astype(np.float64) - ct).max(axis=1) / r["size"][inner]
    print(f"{kind} n={n}: {len(gpu)} nodes; reference vs FP64: mass {rel_m.max():.2e}, "
          f"centre/edge {dc.max():.2e}")
    assert rel_m.max() <= 1e-3 and dc.max() <= 1e-3


def test_reference_traversal_over_export(ctx, orc):
    n, size = 20000, 128.0
    b = ics.disk_bulge(n, seed=31)
    ctx.upload_bodies(b)
    gpu = ctx.bh_tree(tree_params(size))
    ref = orc.bh_build(b, size)
    # FP64 centres recomputed from the leaves: the two trees are the same tree, so the
    # traversal and every force are bit-identical
    g_exp, w_exp = orc.bh_accelerate_f64com(b.copy(), size, theta=0.5, nodes=gpu, visits=True)
    g_ref, w_ref = orc.bh_accelerate_f64com(b.copy(), size, theta=0.5, nodes=ref, visits=True)
    assert np.array_equal(w_exp, w_ref)
    assert np.array_equal(g_exp[:, 8:11].view(np.uint32), g_ref[:, 8:11].view(np.uint32))
    # the reference's FP32 traversal over the exported centres of mass
    a_exp, v_exp = orc.bh_accelerate(b.copy(), size, theta=0.5, nodes=gpu, visits=True)
    same = v_exp == w_ref
    e = np.linalg.norm(a_exp[same, 8:11] - g_ref[same, 8:11], axis=1) / \
        np.linalg.norm(g_ref[same, 8:11], axis=1)
    print(f"identical traversals {same.mean():.5f}; relative error q99.9 "
          f"{np.quantile(e, 0.999):.2e}, max {e.max():.2e}")
    assert same.mean() >= 0.999
    assert np.quantile(e, 0.999) <= 1e-5 and e.max() <= 1e-4


def test_last_build_is_the_pre_step_tree(ctx):
    # Solver::tree() after a step is the tree that step used, built from the positions at
    # its start (source/solvers/gpu_solver_split.h:59-70, 115-118)
    b = ics.plummer(4096, seed=3)
    p = tree_params(64.0, theta=0.5, eps=0.01, dt=1e-2, leaf_size=8)
    ctx.upload_bodies(b)
    before = ctx.bh_tree(p)
    ctx.upload_bodies(b)
    ctx.step(p, 1)
    after = ctx.bh_tree(None)
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
    moved = ctx.bh_tree(p)  # rebuilt from the stepped positions
    assert not np.array_equal(moved["com"], before["com"])
    # a captured step sequence: the export follows the replay's last build
    ctx.step_graph(p, 3)
    ctx.synchronize()
    pre = ctx.download_bodies()
    ctx.step_graph(p, 1)  # a different sequence: captured anew, then replayed
    got = ctx.bh_tree(None)
    ctx.upload_bodies(pre)
    want = ctx.bh_tree(p)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_export_errors_and_empty(ctx):
    c = _lib.Context(0)
    try:
        c.upload_bodies(ics.plummer(64, seed=1))
        with pytest.raises(RuntimeError):
            c.bh_tree(None)  # nothing built yet
        c.reserve(0)
        t = c.bh_tree(tree_params(10.0))
        assert len(t) == 1 and t[0]["mass"] == 0 and t[0]["size"] == 10 and t[0]["children"] == 0
    finally:
        c.close()
