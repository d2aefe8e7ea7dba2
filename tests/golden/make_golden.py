"""Writes tests/golden/golden_small.npz: small seeded inputs with their exact KNN rows.

Expected rows come from tests/np_ref.py (pure numpy: FP64 keys, ties by id, FP32
fixed-order distances in the reference's descending-variance dim order) -- not from the C
oracle, so the fixtures pin the oracle and the GPU path independently.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import np_ref  # noqa: E402


def cases():
    rng = np.random.default_rng(20240601)
    out = {}
    # 1. RA/Dec-like uniform box (the small plain-method workload, scaled down)
    p = np.empty((1500, 2), np.float32)
    p[:, 0] = rng.uniform(0, 360, 1500)
    p[:, 1] = rng.uniform(-90, 90, 1500)
    out["uniform2d_k8"] = (p, 8, 2)
    # 2. clustered 2-D (dense Gaussian blobs on a sparse background)
    u = rng.uniform(0, 50, (700, 2))
    c = rng.uniform(0, 50, (4, 2))[rng.integers(0, 4, 800)] + rng.normal(0, 0.3, (800, 2))
    out["mixed2d_k16"] = (np.concatenate([u, c]).astype(np.float32)[rng.permutation(1500)], 16, 2)
    # 3. 3-D, all dims indexed: the variance order decides the sum order
    q = rng.uniform(0, 1, (1000, 3)) * np.array([1.0, 3.0, 2.0])
    out["uniform3d_k5"] = (q.astype(np.float32), 5, 3)
    # 4. duplicated points: ties broken by id
    base = rng.integers(0, 12, (150, 2)).astype(np.float32)
    out["duplicates2d_k10"] = (np.concatenate([base, base, base, base])[rng.permutation(600)], 10, 2)
    # 5. integer lattice: many exactly equal distances
    g = np.stack(np.meshgrid(np.arange(20), np.arange(20)), -1).reshape(-1, 2).astype(np.float32)
    out["lattice2d_k12"] = (g, 12, 2)
    # 6. N == K+1: every row holds every point
    out["tiny_n_eq_k1"] = (rng.normal(0, 1, (17, 2)).astype(np.float32), 16, 2)
    # 7. 4-D skewed, 2 indexed dims
    s = np.clip(rng.exponential(0.25, (1200, 4)), 0, 1).astype(np.float32)
    out["skewed4d_k32_nd2"] = (s, 32, 2)
    return out


def main():
    arrays = {}
    for name, (pts, k, nd) in cases().items():
        perm, _ = np_ref.reorder_perm(pts)
        pr = np.ascontiguousarray(pts[:, perm])
        ids, dists = np_ref.bruteforce(pr, k)
        arrays[f"{name}__points"] = pts
        arrays[f"{name}__k"] = np.array([k, nd], np.int64)
        arrays[f"{name}__perm"] = perm
        arrays[f"{name}__ids"] = ids
        arrays[f"{name}__dists"] = dists
    path = os.path.join(HERE, "golden_small.npz")
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
