"""Initial conditions drawn on the device (nbhip_generate, nbhip_batch_add_generated),
through the C ABI, against the oracle's restatement (oracle/nbody_oracle.c, orc_generate).

Bar: the integer streams are the same Philox4x64-10 words and both sides evaluate the same
FP64 expressions without FMA contraction, so the FP32 results agree bit for bit except
where the device's and the host's FP64 libm (log, sin, cos, pow, exp, atanh) differ in the
last bit and that reaches an FP32 rounding boundary: >= 99.9% of values bit-exact, every
value within 4 FP32 ulps (relative 4.8e-7; absolute 1e-9 of the model scale near 0).
"""
import numpy as np
import pytest

from nbody_amd import _lib

pytestmark = pytest.mark.gpu

KINDS = {
    "plummer": lambda seed: _lib.ics(_lib.IC_PLUMMER, seed=seed, center=[1, -2, 0.5], vel=[0, 0.1, 0]),
    "disk_bulge": lambda seed: _lib.ics(_lib.IC_DISK_BULGE, seed=seed),
    "disk": lambda seed: _lib.ics(_lib.IC_DISK, seed=seed, scale=100.0),
    "disk_tilted": lambda seed: _lib.ics(_lib.IC_DISK, seed=seed, axis=[0.6, 0.0, 0.8], vel=[3, 0, 0]),
    "cube": lambda seed: _lib.ics(_lib.IC_CUBE, seed=seed, vel=[0, 0, -1]),
}


def compare(gpu, cpu, scale):
    g = gpu.astype(np.float64)
    c = cpu.astype(np.float64)
    diff = np.abs(g - c)
    exact = (gpu.view(np.uint32) == cpu.view(np.uint32)).mean()
    bound = 4.8e-7 * np.abs(c) + 1e-9 * scale
    return exact, bool(np.all(diff <= bound)), float((diff / np.maximum(np.abs(c), 1e-30)).max())


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (7, 3), (1000, 4), (100003, 5)])
def test_generate_matches_oracle(ctx, orc, kind, n, seed):
    spec = KINDS[kind](seed)
    ctx.generate(spec, n)
    assert ctx.count == n
    gpu = ctx.download_bodies()
    cpu = orc.generate(spec, n)
    exact, ok, worst = compare(gpu, cpu, spec.scale)
    print(f"{kind} n={n}: bit-exact {exact:.6f}, worst relative {worst:.2e}")
    assert exact >= 0.999 and ok


def test_generate_full_size_disk_bulge(ctx, orc):
    n = 1 << 20
    spec = _lib.ics(_lib.IC_DISK_BULGE, seed=2026)
    ctx.generate(spec, n)
    gpu = ctx.download_bodies()
    cpu = orc.generate(spec, n)
    exact, ok, worst = compare(gpu, cpu, spec.scale)
    print(f"disk+bulge 2^20: bit-exact {exact:.6f}, worst relative {worst:.2e}")
    assert exact >= 0.999 and ok
    assert np.all(gpu[:, 8:11] == 0)


def test_batch_add_generated_equals_context(ctx):
    specs = [(KINDS["plummer"](11), 3000), (KINDS["cube"](12), 700), (KINDS["disk"](13), 1500)]
    p = _lib.params(dt=1e-3, eps=0.01)
    want = []
    for s, n in specs:
        ctx.generate(s, n)
        want.append(ctx.download_bodies())
    with _lib.Batch(ctx) as batch:
        ids = [batch.add_generated(s, n, p) for s, n in specs]
        got = [batch.download(i) for i in ids]
        # and it steps like the same bodies added from the host
        host = [batch.add(w, p) for w in want]
        batch.step(3)
        for a, b in zip(ids, host):
            assert np.array_equal(batch.download(a).view(np.uint32), batch.download(b).view(np.uint32))
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32))


def test_generate_then_step_equals_upload_then_step(ctx):
    spec = KINDS["disk_tilted"](21)
    p = _lib.params(dt=1e-3, eps=0.0, size=10000.0, G=1.0)
    ctx.generate(spec, 2048)
    b = ctx.download_bodies()
    ctx.step(p, 2)
    a = ctx.download_bodies()
    ctx.upload_bodies(b)
    ctx.step(p, 2)
    assert np.array_equal(a.view(np.uint32), ctx.download_bodies().view(np.uint32))


def test_generate_rejects_bad_specs(ctx):
    bad = _lib.ics(_lib.IC_PLUMMER)
    bad.kind = 9
    with pytest.raises(_lib.HipError):
        ctx.generate(bad, 10)
    with pytest.raises(_lib.HipError):
        ctx.generate(_lib.ics(_lib.IC_PLUMMER, scale=0.0), 10)
    with pytest.raises(_lib.HipError):
        ctx.generate(_lib.ics(_lib.IC_DISK, axis=[0, 0, 0]), 10)
    ctx.generate(_lib.ics(_lib.IC_CUBE), 0)
    assert ctx.count == 0
