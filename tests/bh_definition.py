"""The leaf-bucket octree written from its definition, and a traversal of it.

Shared by the CPU pinning of the bucket-aware oracle (tests/test_oracle.py) and the GPU
node-array check (tests/test_bh_gpu.py).  It is built by a different route than both the
library (scans of sorted-key prefixes on the device) and the oracle (the reference's own
insertion tree, source/bhtree.cpp:14-95, marked with buckets): a recursion over key
prefixes, so an agreement between any two of them is a check, not a restatement.
"""
import numpy as np


def bucket_tree(keys, spm, size, theta, leaf, textbook=False, levels=21):
    """The leaf-bucket octree written from its definition, recursively over key prefixes:
    a cell of two or more bodies is one node at the deepest level its bodies share (chains
    of one-child cells collapse), at most `levels` (21 for one system, 18 in a batch); it is
    a bucket when it holds at most `leaf` bodies or sits at that deepest level, else an
    internal node whose children follow in octant order, a one-body child being a leaf
    entry.  Nodes in pre-order with skip pointers; `spm` holds
    the bodies in key order (x, y, z, mass).  Rows: first, count, next, open_sq, com xyz,
    mass (count 0 = internal node)."""
    K = [int(k) for k in keys]
    n = len(K)
    th2 = np.float32(theta) * np.float32(theta)
    rows = []

    def cpl(i, j):
        x = K[i] ^ K[j]
        return 21 if x == 0 else (64 - x.bit_length() - 1) // 3

    def open_sq(level):
        s = np.float32(size)
        for _ in range(level):
            s = np.float32(s * np.float32(0.5))
        s2 = np.float32(s * s)
        return np.float32(s2 / th2) if textbook else np.float32(s2 * th2)

    def leaf_entry(a):
        # a leaf is always applied: its threshold is -inf
        rows.append([a, 1, len(rows) + 1, np.float32(-np.inf), *spm[a, :3], spm[a, 3]])

    def cell(s, e):
        m = min(cpl(s, e - 1), levels)
        me = len(rows)
        w = spm[s:e, 3].astype(np.float64)
        M = w.sum()
        com = (spm[s:e, :3].astype(np.float64) * w[:, None]).sum(0) / M if M != 0 else np.zeros(3)
        bucket = e - s <= leaf or m == levels
        # a cell of zero mass (as FP32) is never opened: its threshold is -inf
        thr = open_sq(m) if np.float32(M) != 0 else np.float32(-np.inf)
        rows.append([s, e - s if bucket else 0, 0, thr, *com, M])
        if not bucket:
            shift = 60 - 3 * m
            a = s
            while a < e:
                q = (K[a] >> shift) & 7
                b = a
                while b < e and (K[b] >> shift) & 7 == q:
                    b += 1
                if b - a == 1:
                    leaf_entry(a)
                else:
                    cell(a, b)
                a = b
        rows[me][2] = len(rows)

    if n == 1:
        leaf_entry(0)
    else:
        cell(0, n)
    return rows


def walk(rows, spm, radius=None, G=1.0, eps=0.0):
    """Every body's traversal of the rows, all bodies advanced together: a node of zero
    mass is skipped; a leaf entry is applied; another node is applied whole when
    |com - x|^2 > open_sq (FP32, the reference's operation order, com rounded to FP32),
    else an internal node opens to the next row and a bucket is applied body by body.
    Forces in FP64.  Returns (acc (n, 3) in key order, visits)."""
    R = np.array([r[:3] for r in rows], np.int64)
    first, count, nxt = R[:, 0], R[:, 1], R[:, 2]
    thr = np.array([r[3] for r in rows], np.float32)
    com = np.array([r[4:7] for r in rows], np.float64)
    mass = np.array([r[7] for r in rows], np.float64)
    com32 = com.astype(np.float32)
    n = spm.shape[0]
    x32 = spm[:, :3].astype(np.float32)
    x = x32.astype(np.float64)
    ms = spm[:, 3].astype(np.float64)
    r2 = np.zeros(n) if radius is None else np.asarray(radius, np.float64) ** 2
    eps2 = float(eps) ** 2
    total = len(rows)
    idx = np.zeros(n, np.int64)
    acc = np.zeros((n, 3))
    visits = np.zeros(n, np.int64)

    def add(sel, src, m):
        d = src - x[sel]
        d2 = (d * d).sum(1)
        w = np.where(d2 > r2[sel], G * m / np.maximum((d2 + eps2) ** 1.5, 1e-300), 0.0)
        np.add.at(acc, sel, w[:, None] * d)

    while True:
        act = np.nonzero(idx < total)[0]
        if act.size == 0:
            break
        i = idx[act]
        live = mass[i] != 0
        d = com32[i] - x32[act]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        apply = live & ((count[i] == 1) | (d2 > thr[i]))
        opened = live & ~apply
        bucket = opened & (count[i] > 1)
        sel = act[apply]
        add(sel, com[i[apply]], mass[i[apply]])
        visits[sel] += 1
        bsel, bi = act[bucket], i[bucket]
        for t in range(int(count[bi].max()) if bi.size else 0):
            on = t < count[bi]
            src = first[bi[on]] + t
            add(bsel[on], x[src], ms[src])
        visits[bsel] += count[bi]
        idx[act] = np.where(opened & (count[i] == 0), i + 1, nxt[i])
    return acc, visits
