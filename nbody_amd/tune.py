"""Run the direct-sum autotuner on a workload and print every timed launch shape.

    python -m nbody_amd.tune [--n 262144] [--workload two_plummer|plummer] [--eps 0.005]

The shape chosen is kept by the context for that body count (nbhip_tune); this script
only reports the sweep, as a table of (P bodies per lane, W waves per block, j-split --
"pair-symmetric" for the tiles that evaluate each pair once --, ms per force pass,
TFLOP/s at 20 flop per interaction).
"""
from __future__ import annotations

import argparse

from . import _lib, ics


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--workload", choices=["two_plummer", "plummer"], default="two_plummer")
    ap.add_argument("--eps", type=float, default=0.005)
    ap.add_argument("--seed", type=int, default=2026)
    a = ap.parse_args()
    gen = ics.two_plummer if a.workload == "two_plummer" else ics.plummer
    bodies = gen(a.n, seed=a.seed)
    p = _lib.params(eps=a.eps)
    with _lib.Context(0) as ctx:
        ctx.upload_bodies(bodies)
        rows, best = ctx.tune(p)
    inter = float(a.n) ** 2
    print(f"# N={a.n} {a.workload} eps={a.eps}: {len(rows)} shapes")
    print("| P | W | j-split | ms | TFLOP/s |")
    print("|---|---|---|---|---|")
    for P, W, js, ms in sorted(rows, key=lambda r: r[3]):
        print(f"| {P} | {W} | {js if js else 'pair-symmetric'} | {ms:.3f} | {20 * inter / (ms * 1e-3) / 1e12:.1f} |")
    kind = f"j-split={best[2]}" if best[2] else "pair-symmetric"
    print(f"chosen: P={best[0]} W={best[1]} {kind} ({best[3]:.3f} ms)")


if __name__ == "__main__":
    main()
