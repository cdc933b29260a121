"""Write the golden fixtures in tests/golden/ from the reference's own CPU sources.

Runs only in a container that has the reference tree: it builds oracle/_ref/libnbref.so
(the reference's source/bhtree.cpp, vector.cpp, util.cpp plus oracle/ref_shim.cpp) and
records inputs and outputs of the reference functions at small sizes:

  gravity_pairs.npz     detail::gravity on random pairs (source/detail/physics.h:21-42)
  wrap_cases.npz        detail::wrap (source/detail/physics.h:50-57)
  brute_force.npz       the brute-force loop (source/solvers/cpu_brute_force.h:26-40)
                        over util::disk(256) and util::cube(128) (source/util.cpp:14-126)
  integrate.npz         detail::integrate_euler after that (source/detail/physics.h:60-70)
  tree.npz              detail::build_tree node arrays (source/detail/tree.h:14-29)
  barnes_hut.npz        Tree::apply with the gravity callback (source/bhtree.cpp:97-138,
                        source/solvers/cpu_barnes_hut.h:34-42), theta 0.5 and 8
  cell_keys.npz         the Bounds::quadrant descent as 21-level keys
                        (include/nbody/bounds.h:31-52)

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main() -> None:
    O.build(ref=True)
    if O.ref() is None:
        raise SystemExit("oracle/_ref/libnbref.so could not be built (no reference tree?)")
    rng = np.random.default_rng(20261203)

    # ---- single pairs ----
    k = 64
    pos = rng.normal(0, 10, (k, 3)).astype(np.float32)
    src = (pos + rng.normal(0, 3, (k, 3))).astype(np.float32)
    src[:4] = pos[:4]                                          # zero distance
    radius = np.abs(rng.normal(0, 1, k)).astype(np.float32)
    radius[4:8] = 0
    mass = rng.uniform(0.1, 10, k).astype(np.float32)
    G = np.where(np.arange(k) % 2 == 0, 1.0, 4.0).astype(np.float32)
    out = np.stack([O.ref_gravity_pair(pos[i], radius[i], src[i], mass[i], G[i]) for i in range(k)])
    np.savez(os.path.join(HERE, "gravity_pairs.npz"), pos=pos, src=src, radius=radius,
             mass=mass, G=G, acc=out)

    # ---- wrap ----
    x = np.concatenate([rng.uniform(-3000, 3000, 96), [499 + 10, -500, 500, 0, -1e-3, 1e4]]).astype(np.float32)
    size = np.concatenate([np.full(96, 1000.0), [1000, 1000, 1000, 1000, 1000, 0.0]]).astype(np.float32)
    w = np.array([O.ref_wrap(float(a), float(b)) for a, b in zip(x, size)], np.float32)
    np.savez(os.path.join(HERE, "wrap_cases.npz"), x=x, size=size, wrapped=w)

    # ---- brute force + integrate ----
    disk = O.ref_disk(256, 100.0)
    cube = O.ref_cube(128)
    bf_disk = O.ref_brute_force(disk.copy(), 1.0)
    bf_cube = O.ref_brute_force(cube.copy(), 4.0)
    np.savez(os.path.join(HERE, "brute_force.npz"), disk=disk, disk_out=bf_disk,
             cube=cube, cube_out=bf_cube)
    int_wrap = O.ref_integrate(bf_disk.copy(), 1.0 / 60.0, 10000.0, True)
    int_small = O.ref_integrate(bf_disk.copy(), 0.5, 50.0, True)      # many wraps
    int_nowrap = O.ref_integrate(bf_disk.copy(), 0.5, 50.0, False)
    np.savez(os.path.join(HERE, "integrate.npz"), bodies=bf_disk, dt_a=np.float32(1 / 60),
             out_a=int_wrap, out_b=int_small, out_c=int_nowrap)

    # ---- tree ----
    d64 = O.ref_disk(64, 100.0)
    t64 = O.ref_build_tree(d64, 10000.0)
    c100 = O.ref_cube(100)
    t100 = O.ref_build_tree(c100, 1024.0)
    np.savez(os.path.join(HERE, "tree.npz"), disk64=d64, disk64_nodes=t64.view(np.uint8),
             cube100=c100, cube100_nodes=t100.view(np.uint8))

    # ---- Barnes-Hut ----
    d512 = O.ref_disk(512, 100.0)
    bh05, v05 = O.ref_bh_accelerate(d512.copy(), 10000.0, 0.5, 1.0, visits=True)
    bh8, v8 = O.ref_bh_accelerate(d512.copy(), 10000.0, 8.0, 1.0, visits=True)
    np.savez(os.path.join(HERE, "barnes_hut.npz"), bodies=d512, out_theta05=bh05,
             visits_theta05=v05, out_theta8=bh8, visits_theta8=v8)

    # ---- cell keys ----
    d2k = O.ref_disk(2048, 100.0)
    np.savez(os.path.join(HERE, "cell_keys.npz"), bodies=d2k,
             keys_256=O.ref_cell_keys(d2k, 256.0, 21),
             keys_10000=O.ref_cell_keys(d2k, 10000.0, 21),
             keys_100=O.ref_cell_keys(d2k, 100.0, 21))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
