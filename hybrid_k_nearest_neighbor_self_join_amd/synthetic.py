"""Seeded synthetic point sets of the benchmark workloads (numpy PCG64, platform-stable).

  uniform_box        N points, x ~ U[0,360), y ~ U[-90,90]   (RA/Dec-like 2-D box)
  mixed_box          70% uniform in the box + 30% from 64 isotropic Gaussian clusters
                     (sigma 0.5, centres uniform in the box), FP64 draws rounded to FP32
  unit_cube          N points uniform in [0,1]^d
  skewed_cube        each coordinate ~ Exp(lambda=4), clipped to [0,1]
  batched_sets       many independent sets, sizes log-uniform, mixed dims / k
"""
from __future__ import annotations

import numpy as np

BOX_LO = np.array([0.0, -90.0])
BOX_HI = np.array([360.0, 90.0])


def uniform_box(n: int, seed: int = 0x5EED0001) -> np.ndarray:
    rng = np.random.default_rng(seed)
    pts = np.empty((n, 2), np.float32)
    pts[:, 0] = rng.uniform(0.0, 360.0, n).astype(np.float32)
    pts[:, 1] = rng.uniform(-90.0, 90.0, n).astype(np.float32)
    # U[0,360) in FP64 may round up to 360.0f; keep the half-open range
    pts[:, 0] = np.where(pts[:, 0] >= np.float32(360.0), np.nextafter(np.float32(360.0), np.float32(0)), pts[:, 0])
    return pts


def mixed_box(n: int, seed: int = 0x5EED0002, frac_uniform: float = 0.7, nclusters: int = 64,
              sigma: float = 0.5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nu = int(round(n * frac_uniform))
    nc = n - nu
    out = np.empty((n, 2), np.float64)
    out[:nu, 0] = rng.uniform(0.0, 360.0, nu)
    out[:nu, 1] = rng.uniform(-90.0, 90.0, nu)
    centres = np.column_stack([rng.uniform(0.0, 360.0, nclusters), rng.uniform(-90.0, 90.0, nclusters)])
    lab = rng.integers(0, nclusters, nc)
    out[nu:] = centres[lab] + rng.normal(0.0, sigma, (nc, 2))
    perm = rng.permutation(n)
    return np.ascontiguousarray(out[perm].astype(np.float32))


def unit_cube(n: int, d: int, seed: int = 0x5EED0004) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, (n, d)).astype(np.float32)


def skewed_cube(n: int, d: int, seed: int = 0x5EED0005, lam: float = 4.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.clip(rng.exponential(1.0 / lam, (n, d)), 0.0, 1.0).astype(np.float32)


def batched_sets(nsets: int = 256, seed: int = 0x5EED0006, lo: int = 1000, hi: int = 200000,
                 dims=(2, 3, 4, 6), ks=(8, 32)):
    """Many independent sets: sizes log-uniform in [lo, hi], d and k drawn from `dims` / `ks`;
    each set is 70% uniform in [0,1]^d plus 30% from 2..16 Gaussian clusters (sigma 0.01,
    centres uniform), the mix of the 10M mixed_box set.  Returns a list of (points, k)."""
    rng = np.random.default_rng(seed)
    sets = []
    for _ in range(nsets):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        d = int(rng.choice(dims))
        k = int(rng.choice(ks))
        nu = int(round(0.7 * n))
        nc = int(rng.integers(2, 17))
        centres = rng.uniform(0.0, 1.0, (nc, d))
        pts = np.concatenate([rng.uniform(0.0, 1.0, (nu, d)),
                              centres[rng.integers(0, nc, n - nu)] + rng.normal(0.0, 0.01, (n - nu, d))])
        pts = pts[rng.permutation(n)]
        sets.append((np.ascontiguousarray(pts.astype(np.float32)), k))
    return sets
