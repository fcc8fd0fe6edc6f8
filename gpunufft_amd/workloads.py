"""Seeded synthetic MRI workloads: trajectories (SoA float32 in [-0.5, 0.5)),
density weights, coil sensitivities and complex data.

  radial_2d       spokes x samples, angles uniform over pi (or golden angle)
  koosh_3d        Fibonacci-sphere spokes, radii linear in [-0.5, 0.5)
  spiral_2d       variable-density spiral interleaves
  gaussian_sens   smooth complex Gaussian-blob coil maps, sum |S_c|^2 = 1
"""

from __future__ import annotations

import math

import numpy as np

GOLDEN_ANGLE = math.pi / ((1 + math.sqrt(5)) / 2)


def complex_normal(rng, shape):
    return ((rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) /
            math.sqrt(2)).astype(np.complex64)


def radial_2d(n_spokes, n_samples, golden=False):
    r = np.linspace(-0.5, 0.5, n_samples, endpoint=False)
    if golden:
        th = np.arange(n_spokes) * GOLDEN_ANGLE
    else:
        th = math.pi * np.arange(n_spokes) / n_spokes
    kx = np.outer(np.cos(th), r).ravel()
    ky = np.outer(np.sin(th), r).ravel()
    return np.stack([kx, ky]).astype(np.float32)


def koosh_3d(n_spokes, n_samples):
    i = np.arange(n_spokes)
    z = 1 - 2 * (i + 0.5) / n_spokes
    phi = i * math.pi * (3 - math.sqrt(5))
    rho = np.sqrt(np.maximum(0.0, 1 - z * z))
    d = np.stack([rho * np.cos(phi), rho * np.sin(phi), z])
    r = np.linspace(-0.5, 0.5, n_samples, endpoint=False)
    return np.stack([np.outer(d[a], r).ravel() for a in range(3)]).astype(np.float32)


def spiral_2d(n_interleaves, n_samples, n_turns=16, alpha=1.5):
    t = np.linspace(0, 1, n_samples, endpoint=False)
    ta = t ** alpha
    ks = []
    for l in range(n_interleaves):
        ph = 2 * math.pi * n_turns * ta + 2 * math.pi * l / n_interleaves
        ks.append(0.5 * ta * np.exp(1j * ph))
    k = np.concatenate(ks)
    k = np.clip(k.real, -0.5, 0.4999999) + 1j * np.clip(k.imag, -0.5, 0.4999999)
    return np.stack([k.real, k.imag]).astype(np.float32)


def ramp_density(traj, power=1):
    r = np.sqrt(np.sum(traj.astype(np.float64) ** 2, axis=0)) ** power
    m = r.max()
    return (r / m if m > 0 else np.ones_like(r)).astype(np.float32)


def gaussian_sens(img_dims, n_coils, seed, sigma=0.3, ring=0.35):
    """img_dims (x, y[, z]); returns complex64 (C, [z,] y, x)."""
    rng = np.random.default_rng(seed)
    dims = list(img_dims)
    axes = [np.linspace(-0.5, 0.5, n, endpoint=False) for n in dims]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]  # x, y[, z] arrays in [z,]y,x order
    maps = []
    for c in range(n_coils):
        ang = rng.uniform(0, 2 * math.pi)
        centre = [ring * math.cos(ang), ring * math.sin(ang)]
        if len(dims) == 3:
            centre.append(rng.uniform(-ring, ring))
        d2 = sum((g - c0) ** 2 for g, c0 in zip(grids, centre))
        ph = rng.uniform(0, 2 * math.pi)
        maps.append(np.exp(-d2 / (2 * sigma ** 2)) * np.exp(1j * ph))
    s = np.stack(maps)
    s /= np.sqrt(np.sum(np.abs(s) ** 2, axis=0, keepdims=True))
    return s.astype(np.complex64)


BATCH_SIZES = (96, 128, 192, 256, 320, 384)


def uneven_batch_2d(n_ops=64, seed0=100):
    """Batched uneven 2-D workload: n_ops independent operators, operator j
    seeded with seed0 + j. Image N uniform over BATCH_SIZES; variable-density
    spiral k(t) = 0.5 t^1.5 exp(i(2 pi n_turns t^1.5 + 2 pi l / L)) with
    L in [8, 48] interleaves of [1024, 4096] samples and n_turns = N / (2 L)
    (one interleave-turn spacing of L/N in k); coils in [4, 16]; osf 2,
    kernel width 3, |k| density, Gaussian-blob coil maps.
    Returns a list of dicts (traj, img_dims, coils, dens, sens, osf, kw, seed)."""
    ops = []
    for j in range(n_ops):
        rng = np.random.default_rng(seed0 + j)
        N = int(rng.choice(BATCH_SIZES))
        L = int(rng.integers(8, 49))
        ns = int(rng.integers(1024, 4097))
        C = int(rng.integers(4, 17))
        traj = spiral_2d(L, ns, n_turns=max(1.0, N / (2.0 * L)))
        ops.append(dict(traj=traj, img_dims=(N, N), coils=C, dens=ramp_density(traj),
                        sens=gaussian_sens((N, N), C, seed=seed0 + j), osf=2.0, kw=3,
                        seed=seed0 + j))
    return ops


def config(name):
    """Workload of a named configuration: dict(traj, img_dims, osf, kw, sw, coils, dens, seed)."""
    if name == "radial2d_64":
        traj = radial_2d(64, 128)
        return dict(traj=traj, img_dims=(64, 64), osf=2.0, kw=3, sw=8, coils=1,
                    dens=ramp_density(traj), seed=0)
    if name == "radial2d_256_mc16":
        traj = radial_2d(402, 512, golden=True)
        return dict(traj=traj, img_dims=(256, 256), osf=2.0, kw=3, sw=8, coils=16,
                    dens=ramp_density(traj), seed=1)
    if name == "koosh3d_128_mc8":
        traj = koosh_3d(16384, 256)
        return dict(traj=traj, img_dims=(128, 128, 128), osf=1.5, kw=3, sw=8, coils=8,
                    dens=ramp_density(traj, 2), seed=2)
    raise KeyError(name)
