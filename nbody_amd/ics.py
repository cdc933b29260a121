"""Seeded initial conditions for the benchmark and test workloads.

The reference ships only `util::disk` and `util::cube` (reference source/util.cpp:14-126),
which draw from an unseeded std::default_random_engine.  The workloads here need seeded
Plummer spheres, a two-sphere collision and an exponential disk with a Plummer bulge, so
they are generated here with numpy's PCG64 stream for the given seed.

Every generator returns reference-layout bodies: float32 array (n, 12) with
pos[0:3], radius[3], vel[4:7], mass[7], acc[8:11], pad[11] (include/nbody/body.h:6-19).
"""
from __future__ import annotations

import numpy as np

# Plummer spheres are truncated at this enclosed-mass fraction (r ~ 22 a), the usual
# practice so one stray body does not set the size of the system.
PLUMMER_MASS_CUT = 0.999


def empty(n: int) -> np.ndarray:
    return np.zeros((n, 12), np.float32)


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    cos_t = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sin_t = np.sqrt(1.0 - cos_t * cos_t)
    return np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], axis=1)


def plummer_f64(n: int, seed: int, a: float = 1.0, mass: float = 1.0, G: float = 1.0):
    """Positions and velocities (float64) of a Plummer sphere in its own centre-of-mass
    frame, by the Aarseth, Henon & Wielen (1974) method."""
    rng = np.random.default_rng(seed)
    x1 = rng.uniform(0.0, PLUMMER_MASS_CUT, n)
    x1 = np.maximum(x1, 1e-300)
    r = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    pos = r[:, None] * _unit_vectors(rng, n)
    # speed: q = v / v_esc from g(q) = q^2 (1 - q^2)^{7/2} by rejection (max < 0.1)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.uniform(0.0, 1.0, todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 * x4 * (1.0 - x4 * x4) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    v_esc = np.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    vel = (q * v_esc)[:, None] * _unit_vectors(rng, n)
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    pos *= a
    vel *= np.sqrt(G * mass / a)
    return pos, vel


def _pack(pos, vel, m, radius: float = 0.0) -> np.ndarray:
    b = empty(pos.shape[0])
    b[:, 0:3] = pos
    b[:, 3] = radius
    b[:, 4:7] = vel
    b[:, 7] = m
    return b


def plummer(n: int, seed: int, a: float = 1.0, mass: float = 1.0, G: float = 1.0) -> np.ndarray:
    """One Plummer sphere: scale radius a, total mass `mass`, equal masses mass/n."""
    pos, vel = plummer_f64(n, seed, a, mass, G)
    return _pack(pos, vel, mass / n)


def two_plummer(n: int, seed: int, sep: float = 4.0, speed: float = 0.3, a: float = 1.0,
                mass: float = 1.0, G: float = 1.0) -> np.ndarray:
    """Two Plummer spheres of n/2 bodies and mass/2 each at (+-sep, 0, 0), moving towards
    each other at (-+speed, 0, 0)."""
    h = n // 2
    p1, v1 = plummer_f64(h, seed, a, mass / 2, G)
    p2, v2 = plummer_f64(n - h, seed + 0x9E3779B9, a, mass / 2, G)
    p1[:, 0] += sep
    v1[:, 0] -= speed
    p2[:, 0] -= sep
    v2[:, 0] += speed
    return _pack(np.concatenate([p1, p2]), np.concatenate([v1, v2]), mass / n)


def disk_bulge(n: int, seed: int, rd: float = 3.0, z0: float = 0.3, disk_frac: float = 0.8,
               bulge_a: float = 0.5, mass: float = 1.0, G: float = 1.0,
               rmax_scale: float = 10.0) -> np.ndarray:
    """Exponential disk (scale length rd, sech^2 scale height z0, `disk_frac` of the mass,
    circular velocities from the enclosed mass) plus a Plummer bulge (scale radius bulge_a,
    the rest of the mass, Plummer velocities)."""
    rng = np.random.default_rng(seed)
    nd = int(round(n * disk_frac))
    nb = n - nd
    # disk: R ~ Gamma(2, rd) truncated at rmax_scale * rd; z ~ z0 atanh(2u - 1)
    R = np.empty(nd)
    todo = np.arange(nd)
    while todo.size:
        s = rng.gamma(2.0, rd, todo.size)
        ok = s < rmax_scale * rd
        R[todo[ok]] = s[ok]
        todo = todo[~ok]
    phi = rng.uniform(0.0, 2.0 * np.pi, nd)
    u = rng.uniform(1e-12, 1.0 - 1e-12, nd)
    z = z0 * np.arctanh(2.0 * u - 1.0)
    dpos = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    bpos, bvel = plummer_f64(nb, seed + 0x51ED270B, bulge_a, mass * (1.0 - disk_frac), G) \
        if nb else (np.zeros((0, 3)), np.zeros((0, 3)))
    m = mass / n
    # circular speed from the mass enclosed within each disk body's radius (spherical
    # approximation over all bodies)
    r_all = np.sqrt(np.concatenate([(dpos ** 2).sum(1), (bpos ** 2).sum(1)]))
    r_sorted = np.sort(r_all)
    r_disk = np.sqrt((dpos ** 2).sum(1))
    m_enc = np.searchsorted(r_sorted, r_disk, side="right") * m
    vc = np.sqrt(G * m_enc / np.maximum(r_disk, 1e-12))
    dvel = np.stack([-vc * np.sin(phi), vc * np.cos(phi), np.zeros(nd)], axis=1)
    pos = np.concatenate([dpos, bpos])
    vel = np.concatenate([dvel, bvel])
    pos -= pos.mean(axis=0)
    vel -= vel.mean(axis=0)
    return _pack(pos, vel, m)


def cube(n: int, seed: int, size: float = 500.0, mass: float = 1.0) -> np.ndarray:
    """Uniform cube of edge `size` centred on the origin, at rest, masses `mass` each."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5 * size, 0.5 * size, (n, 3))
    return _pack(pos, np.zeros((n, 3)), mass)


def split(bodies: np.ndarray):
    """Reference AoS bodies -> (pos_mass, vel_radius, acc) float4 arrays (source/gpu.h:32-48)."""
    b = np.ascontiguousarray(bodies, np.float32)
    pm = np.ascontiguousarray(b[:, [0, 1, 2, 7]])
    vr = np.ascontiguousarray(b[:, [4, 5, 6, 3]])
    ac = np.zeros((b.shape[0], 4), np.float32)
    ac[:, 0:3] = b[:, 8:11]
    return pm, vr, ac
