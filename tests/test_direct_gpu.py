"""Parity of the HIP direct-sum path with the CPU oracle, through the C ABI.

Tolerances (all stated here, per body, relative L2 norm |a - a_ref| / |a_ref|):
  * vs the FP64 direct sum:   <= 1e-5 for every checked body (the product target);
  * vs the reference FP32 CPU loop: <= 1e-5 for 99.9% of bodies and <= 1e-3 max -- the
    reference sums 2^12..2^18 FP32 terms sequentially, so its own error (measured against
    FP64) exceeds 1e-5 for bodies whose force nearly cancels; the FP64 check is the
    sharper one;
  * integrate / wrap: bit-exact (same operations in the same order, no FMA contraction).
"""
import os

import numpy as np
import pytest

from nbody_amd import _lib, ics

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def rel_err(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.linalg.norm(a - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)


def accelerate(ctx, bodies, **kw):
    ctx.upload_bodies(bodies)
    p = _lib.params(**kw)
    ctx.accelerate(p)
    return ctx.download_bodies()


def test_golden_brute_force_disk(ctx):
    # util::disk(256) with radius cutoffs and a 4.1e6 central mass; reference output
    g = np.load(os.path.join(GOLD, "brute_force.npz"))
    out = accelerate(ctx, g["disk"], G=1.0)
    ref = g["disk_out"][:, 8:11]
    e = rel_err(out[:, 8:11], ref)
    print("golden disk: max rel", e.max(), "median", np.median(e))
    assert np.quantile(e, 0.999) <= 1e-5 and e.max() <= 1e-3
    out = accelerate(ctx, g["cube"], G=4.0)
    e = rel_err(out[:, 8:11], g["cube_out"][:, 8:11])
    print("golden cube: max rel", e.max())
    assert np.quantile(e, 0.999) <= 1e-5 and e.max() <= 1e-3


@pytest.mark.parametrize("n,eps", [(1, 0.01), (2, 0.0), (3, 0.01), (255, 0.01), (1000, 0.0),
                                   (4096, 0.01), (5000, 0.005)])
def test_direct_vs_f64(ctx, orc, n, eps):
    b = ics.plummer(n, seed=100 + n)
    out = accelerate(ctx, b, eps=eps)
    a64 = orc.direct_f64(b, eps=eps)
    if n == 1:
        assert np.all(out[:, 8:11] == 0)
        return
    e = rel_err(out[:, 8:11], a64)
    print(f"n={n} eps={eps}: max rel vs f64 {e.max():.3e} median {np.median(e):.3e}")
    assert e.max() <= 1e-5
    cpu = b.copy()
    orc.brute_force(cpu, eps=eps)
    e2 = rel_err(out[:, 8:11], cpu[:, 8:11])
    assert np.quantile(e2, 0.999) <= 1e-5 and e2.max() <= 1e-3


def test_integrate_bit_exact(ctx, orc):
    g = np.load(os.path.join(GOLD, "integrate.npz"))
    for out, dt, size, wrap in [("out_a", 1 / 60, 10000.0, 1), ("out_b", 0.5, 50.0, 1),
                                ("out_c", 0.5, 50.0, 0)]:
        ctx.upload_bodies(g["bodies"])
        ctx.integrate(_lib.params(dt=np.float32(dt), size=size, wrap=wrap))
        res = ctx.download_bodies()
        assert np.array_equal(bits(res), bits(g[out])), out


def test_wrap_reference_case(ctx):
    # tests/source/test_wrap.cpp:15-56: x=499, v=10, dt=1, size=1000, G=0
    b = np.zeros((1, 12), np.float32)
    b[0, 0:3] = [499, 0, 0]
    b[0, 3] = 1.0
    b[0, 4:7] = [10, 0, 0]
    b[0, 7] = 1.0
    ctx.upload_bodies(b)
    ctx.step(_lib.params(dt=1.0, size=1000.0, G=0.0))
    x = ctx.download_bodies()
    assert x[0, 0] == -491.0 and abs(x[0, 1]) <= 500 and abs(x[0, 2]) <= 500
    ctx.upload_bodies(b)
    ctx.step(_lib.params(dt=1.0, size=1000.0, G=0.0, wrap=0))
    assert ctx.download_bodies()[0, 0] == 509.0


def test_fused_step_equals_accelerate_then_integrate(ctx):
    b = ics.two_plummer(3000, seed=9)
    p = _lib.params(dt=1e-3, eps=0.01, wrap=1, size=64.0)
    ctx.upload_bodies(b)
    ctx.step(p, 3)
    fused = ctx.download_bodies()
    ctx.upload_bodies(b)
    for _ in range(3):
        ctx.accelerate(p)
        ctx.integrate(p)
    split = ctx.download_bodies()
    assert np.array_equal(bits(fused), bits(split))


def test_steps_match_oracle(ctx, orc):
    # k steps on the GPU vs k steps of the CPU oracle from the same seeded input
    n, k = 2048, 10
    b = ics.plummer(n, seed=31)
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0, wrap=1)
    ctx.upload_bodies(b)
    ctx.step(p, k)
    gpu = ctx.download_bodies()
    cpu = b.copy()
    for _ in range(k):
        orc.brute_force(cpu, eps=0.01)
        orc.integrate(cpu, np.float32(1e-3), 10000.0, True)
    dx = np.abs(gpu[:, 0:3] - cpu[:, 0:3]).max()
    dv = rel_err(gpu[:, 4:7], cpu[:, 4:7])
    print("after", k, "steps: max |dx|", dx, "max rel dv", dv.max())
    assert dx <= 1e-5 and np.quantile(dv, 0.999) <= 1e-5


def test_gravity_scaling_and_momentum(ctx):
    b = ics.plummer(4000, seed=5)
    a1 = accelerate(ctx, b, eps=0.01, G=1.0)[:, 8:11].astype(np.float64)
    a4 = accelerate(ctx, b, eps=0.01, G=4.0)[:, 8:11].astype(np.float64)
    assert np.allclose(a4, 4 * a1, rtol=1e-6, atol=0)
    m = b[:, 7:8].astype(np.float64)
    ma = (m * a1)
    assert np.abs(ma.sum(0)).max() < 1e-5 * np.abs(ma).sum(0).max()


def test_force_error_probe_matches_oracle(ctx, orc):
    b = ics.two_plummer(6000, seed=12)
    ctx.upload_bodies(b)
    p = _lib.params(eps=0.005)
    ctx.accelerate(p)
    s = np.random.default_rng(1).choice(6000, 64, replace=False).astype(np.uint32)
    acc64, rel = ctx.force_error_f64(p, s)
    ref = orc.direct_f64(b, s, eps=0.005)
    # both FP64; only the summation order differs (device tree reduction vs sequential)
    # (bodies whose force nearly cancels reach ~2e-9); the probe serves 1e-5-level checks
    assert rel_err(acc64, ref).max() <= 1e-8
    assert rel.max() <= 1e-5


def test_diagnostics_match_oracle(ctx, orc):
    b = ics.two_plummer(3000, seed=4)
    ctx.upload_bodies(b)
    d = ctx.diagnostics(_lib.params(eps=0.01))
    e = orc.energy_f64(b, eps=0.01)
    assert np.isclose(d[0], e[0], rtol=1e-12)
    assert np.isclose(d[1], e[1], rtol=1e-6)     # potential summed in FP32 per body
    assert np.allclose(d[2:8], e[2:8], rtol=1e-9, atol=1e-12)
    assert d[8] == pytest.approx(e[8], rel=1e-12)


def test_empty_and_shrink(ctx):
    ctx.reserve(0)
    ctx.step(_lib.params(dt=0.1), 2)        # N = 0 is a no-op
    assert ctx.count == 0
    b = ics.cube(4096, seed=3)
    ctx.upload_bodies(b)
    ctx.step(_lib.params(dt=1 / 60, eps=0.0), 1)
    ctx.upload_bodies(b[:16])
    ctx.step(_lib.params(dt=1 / 60, eps=0.0), 1)
    assert ctx.download_bodies().shape == (16, 12)


@pytest.mark.parametrize("P,W,JS", [(1, 4, 0), (2, 2, 0), (8, 4, 0), (4, 8, 0), (4, 4, 3), (8, 1, 1)])
def test_launch_configs_agree(ctx, monkeypatch, P, W, JS):
    # Every compiled (P, W) shape and j-split gives the same forces to the FP32 tolerance
    b = ics.plummer(5000, seed=77)
    ctx.upload_bodies(b)
    p = _lib.params(eps=0.01)
    ctx.accelerate(p)
    base = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3]
    monkeypatch.setenv("NBHIP_DIRECT_P", str(P))
    monkeypatch.setenv("NBHIP_DIRECT_W", str(W))
    monkeypatch.setenv("NBHIP_DIRECT_JSPLIT", str(JS))
    ctx.accelerate(p)
    other = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3]
    assert rel_err(other, base).max() <= 2e-6


def test_full_size_c2_properties(ctx):
    # N = 262,144 two-Plummer collision: momentum conservation and FP64 sampled error
    n = 262144
    b = ics.two_plummer(n, seed=2026)
    ctx.upload_bodies(b)
    p = _lib.params(dt=5e-4, eps=0.005, size=10000.0, wrap=1)
    ctx.accelerate(p)
    s = np.random.default_rng(2026).choice(n, 1024, replace=False).astype(np.uint32)
    _, rel = ctx.force_error_f64(p, s)
    print("N=262144 rel err vs f64: max", rel.max(), "median", np.median(rel))
    assert rel.max() <= 1e-5
    acc = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3].astype(np.float64)
    ma = acc * b[:, 7:8]
    assert np.abs(ma.sum(0)).max() < 1e-5 * np.abs(ma).sum(0).max()


def test_autotuner_picks_a_shape_and_keeps_results(ctx, orc):
    b = ics.plummer(20000, seed=21)
    ctx.upload_bodies(b)
    p = _lib.params(eps=0.01)
    rows, best = ctx.tune(p)
    assert len(rows) >= 10 and best[3] == min(r[3] for r in rows)
    ctx.accelerate(p)                      # now with the tuned shape
    a = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3]
    s = np.arange(0, 20000, 101, dtype=np.uint32)
    assert rel_err(a[s], orc.direct_f64(b, s, eps=0.01)).max() <= 1e-5


@pytest.mark.parametrize("mode", [_lib.MODE_N2, _lib.MODE_NLOGN])
def test_step_graph_replays_equal_eager(ctx, mode):
    # nbhip_step_graph replays a captured hipGraph; every replay (with eager work and
    # synchronisations in between) must equal the eager path bit for bit
    b = ics.disk_bulge(6000, seed=41)
    p = _lib.params(mode=mode, dt=1e-3, eps=0.01, size=128.0, wrap=1, theta=0.5, leaf_size=8)
    results = {}
    for graph in [False, True]:
        ctx.upload_bodies(b)
        snaps = []
        for _ in range(3):
            if graph:
                ctx.step_graph(p, 4)       # replayed from the cached graph after the first
            else:
                ctx.step(p, 4)
            snaps.append(ctx.download_bodies())
            ctx.accelerate(p)              # eager work between replays
            ctx.synchronize()
        results[graph] = snaps
    for e, g in zip(results[False], results[True]):
        assert np.array_equal(bits(e), bits(g))


def test_c1_plain_direct_sum_100_steps(ctx, orc):
    # The reference's CPU-runnable case: N=4096 Plummer (a=1, M=1, G=1), eps=0.01, dt=1e-3,
    # wrap off, 100 steps, GPU vs the reference brute-force restatement stepping alongside.
    # Two FP32 trajectories that differ only in force rounding (the reference loop vs FP64
    # forces rounded to FP32) end 2.4e-7 apart in position and 3.3e-7 (relative) in velocity
    # on this input, so the bars are positions within 1e-5 (absolute, the cloud has radius
    # ~1), velocities within 1e-5 relative, and energy drifts equal to 1e-6.
    n, steps = 4096, 100
    b = ics.plummer(n, seed=2026)
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0, wrap=0)
    ctx.upload_bodies(b)
    e0 = ctx.diagnostics(p)
    for _ in range(steps):
        ctx.step(p, 1)
    gpu = ctx.download_bodies()
    e1 = ctx.diagnostics(p)
    cpu = b.copy()
    for _ in range(steps):
        orc.brute_force(cpu, eps=0.01)
        orc.integrate(cpu, np.float32(1e-3), 10000.0, False)
    dx = np.abs(gpu[:, 0:3] - cpu[:, 0:3]).max()
    dv = rel_err(gpu[:, 4:7], cpu[:, 4:7])
    ec0 = orc.energy_f64(b, eps=0.01)
    ec1 = orc.energy_f64(cpu, eps=0.01)
    drift_gpu = ((e1[0] + e1[1]) - (e0[0] + e0[1])) / abs(e0[0] + e0[1])
    drift_cpu = ((ec1[0] + ec1[1]) - (ec0[0] + ec0[1])) / abs(ec0[0] + ec0[1])
    print(f"C1 after {steps} steps: max |dx| {dx:.2e}, max rel dv {dv.max():.2e}, "
          f"energy drift gpu {drift_gpu:.3e} cpu {drift_cpu:.3e}")
    assert dx <= 1e-5 and dv.max() <= 1e-5
    assert abs(drift_gpu - drift_cpu) <= 1e-6


def test_predicted_readback_is_current_or_refreshed(ctx):
    # nbhip_readback stages the bodies behind the enqueued work; nbhip_collect_bodies uses
    # that copy only while nothing has changed the bodies since, else downloads afresh
    b = ics.plummer(3000, seed=41)
    p = _lib.params(dt=1e-3, eps=0.01)
    ctx.upload_bodies(b)
    ctx.step(p, 2)
    ctx.readback()
    got, staged = ctx.collect_bodies()
    assert staged and np.array_equal(bits(got), bits(ctx.download_bodies()))
    ctx.readback()
    ctx.step(p, 1)                       # a dispatch after the readback: stale
    got, staged = ctx.collect_bodies()
    assert not staged and np.array_equal(bits(got), bits(ctx.download_bodies()))
    ctx.readback()
    ctx.upload_bodies(b)                 # an upload after the readback: stale
    got, staged = ctx.collect_bodies()
    assert not staged and np.array_equal(bits(got), bits(b))
    ctx.upload_bodies(ics.plummer(70000, seed=42))   # the staging memory grows
    ctx.step(p, 1)
    ctx.readback()
    got, staged = ctx.collect_bodies()
    assert staged and got.shape == (70000, 12)
    assert np.array_equal(bits(got), bits(ctx.download_bodies()))
    ctx.reserve(0)
    ctx.readback()
    got, staged = ctx.collect_bodies()
    assert staged and got.shape == (0, 12)


# ---- the pair-symmetric pass (each pair evaluated once, forces both ways) -------------------

@pytest.mark.parametrize("P,W", [(2, 4), (4, 4), (8, 4), (4, 2), (8, 2)])
@pytest.mark.parametrize("n,eps,radius", [(5000, 0.01, 0.0), (2047, 0.0, 0.0), (9001, 0.02, 0.01),
                                          (40000, 0.005, 0.0), (513, 0.01, 0.0)])
def test_symmetric_pass_vs_f64(ctx, orc, monkeypatch, P, W, n, eps, radius):
    # Ragged block counts (n not a multiple of the 64*P*W block), a single block (diagonal
    # tile only), the reference law (eps = 0, cut at zero distance) and per-body radius
    # cutoffs, which apply to the receiving body on each side of a pair: every body within
    # 1e-5 of FP64, and within 2e-6 of the tiled one-sided pass
    b = ics.plummer(n, seed=n + P * 10 + W)
    if radius:
        b[:, 3] = np.random.default_rng(n).uniform(0.0, radius, n).astype(np.float32)
    p = _lib.params(eps=eps, G=2.0)
    ctx.upload_bodies(b)
    ctx.accelerate(p)
    tiled = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3]
    monkeypatch.setenv("NBHIP_DIRECT_SYM", "1")
    monkeypatch.setenv("NBHIP_DIRECT_P", str(P))
    monkeypatch.setenv("NBHIP_DIRECT_W", str(W))
    ctx.upload_bodies(b)
    ctx.accelerate(p)
    sym = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3]
    a64 = orc.direct_f64(b, G=2.0, eps=eps)
    e = rel_err(sym, a64)
    print(f"symmetric P{P} W{W} n={n}: max {e.max():.2e} vs tiled {rel_err(sym, tiled).max():.2e}")
    assert e.max() <= 1e-5
    assert rel_err(sym, tiled).max() <= 2e-6


def test_symmetric_steps_graph_and_momentum(ctx, orc, monkeypatch):
    # fused steps through the symmetric pass: replayed from a graph equals eager bit for bit,
    # 10 steps match 10 oracle steps, and sum m a = 0 holds to rounding (each pair's two
    # forces are one computed value and its negation)
    monkeypatch.setenv("NBHIP_DIRECT_SYM", "1")
    b = ics.plummer(20000, seed=91)
    p = _lib.params(dt=1e-3, eps=0.01, size=10000.0, wrap=1)
    ctx.upload_bodies(b)
    ctx.step(p, 5)
    eager = ctx.download_bodies()
    ctx.upload_bodies(b)
    ctx.step_graph(p, 5)
    ctx.accelerate(p)
    ctx.upload_bodies(b)
    ctx.step_graph(p, 5)
    graph = ctx.download_bodies()
    assert np.array_equal(graph.view(np.uint32), eager.view(np.uint32))
    ctx.upload_bodies(b)
    ctx.step(p, 10)
    gpu = ctx.download_bodies()
    cpu = b.copy()
    for _ in range(10):
        orc.brute_force(cpu, eps=0.01)
        orc.integrate(cpu, np.float32(1e-3), 10000.0, True)
    assert np.abs(gpu[:, 0:3] - cpu[:, 0:3]).max() <= 1e-5
    ctx.upload_bodies(b)
    ctx.accelerate(_lib.params(eps=0.01))
    acc = ctx.download(_lib.RB_ACCELERATIONS)[2][:, :3].astype(np.float64)
    ma = acc * b[:, 7:8]
    assert np.abs(ma.sum(0)).max() < 1e-5 * np.abs(ma).sum(0).max()


def test_symmetric_full_size_c2(ctx, monkeypatch):
    # N = 262,144 two-Plummer collision through the symmetric pass: 1024 sampled bodies
    # within 1e-5 of FP64
    monkeypatch.setenv("NBHIP_DIRECT_SYM", "1")
    n = 262144
    b = ics.two_plummer(n, seed=2026)
    ctx.upload_bodies(b)
    p = _lib.params(dt=5e-4, eps=0.005, size=10000.0, wrap=1)
    ctx.accelerate(p)
    s = np.random.default_rng(2026).choice(n, 1024, replace=False).astype(np.uint32)
    _, rel = ctx.force_error_f64(p, s)
    print("symmetric N=262144 rel err vs f64: max", rel.max(), "median", np.median(rel))
    assert rel.max() <= 1e-5
