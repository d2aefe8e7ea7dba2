"""Learned per-query initial radius: training data, training, calibration and model files.

The network (MLP 12 -> 64 -> 64 -> 1, ReLU, FP32) predicts log(r_K1 / wf): the distance to a
query's (K+1)-th nearest point relative to the feature-grid width wf.  Training targets come
from exact joins computed by this library (plain schedule) on seeded synthetic sets; features
and training run on the GPU through the C ABI (knnjoin_radius_features, knnjoin_mlp_train).
The model only sets the radius a query is first tried at; rows stay exact whatever it predicts.

    python -m hybrid_k_nearest_neighbor_self_join_amd.learned_radius train [--d 2] [--nd 2] [--out DIR]

A model is specific to the number of indexed dims `nd` and, for data with more dims than are
indexed, to NDIM `d`: files radius_mlp_nd{nd}.npz (2-D data, registered for any NDIM) and
radius_mlp_d{d}_nd{nd}.npz (registered for NDIM = d only).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import time

import numpy as np

from . import _lib, synthetic

PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(PKG, "weights")
MLP_NW = _lib.MLP_NW


def _stem(nd: int, d: int) -> str:
    return f"radius_mlp_nd{nd}" if d == nd == 2 else f"radius_mlp_d{d}_nd{nd}"


def weights_path(nd: int, directory: str = WEIGHTS_DIR, d: int | None = None) -> str:
    return os.path.join(directory, _stem(nd, nd if d is None else d) + ".npz")


def raw_path(nd: int, directory: str = WEIGHTS_DIR, d: int | None = None) -> str:
    """Model file of the C++ command line (`knnjoin_main --learned FILE`): the 5057 weights
    then the margin, little-endian float32."""
    return os.path.join(directory, _stem(nd, nd if d is None else d) + ".bin")


def save(nd: int, w: np.ndarray, margin: float, info: dict, directory: str = WEIGHTS_DIR,
         d: int | None = None):
    os.makedirs(directory, exist_ok=True)
    np.savez(weights_path(nd, directory, d), weights=w.astype(np.float32), margin=np.float32(margin),
             info=np.array(json.dumps(info)))
    with open(raw_path(nd, directory, d), "wb") as f:
        f.write(np.concatenate([w.astype("<f4"), np.array([margin], "<f4")]).tobytes())


def init_weights(seed: int) -> np.ndarray:
    """He-normal weights, zero biases, in the flat layout W1, b1, W2, b2, w3, b3."""
    rng = np.random.default_rng(seed)
    w = np.zeros(MLP_NW, np.float32)
    w[0:768] = rng.normal(0.0, np.sqrt(2.0 / 12), 768)
    w[832:4928] = rng.normal(0.0, np.sqrt(2.0 / 64), 4096)
    w[4992:5056] = rng.normal(0.0, np.sqrt(1.0 / 64), 64)
    return w


def training_pairs(points: np.ndarray, k: int, nd: int, max_queries: int, seed: int):
    """(features, targets) of a random subset of queries: targets = log(d_K / wf), d_K the
    exact (K+1)-th neighbour distance from a plain-schedule join of `points`."""
    ids, dists, st = _lib.knnjoin_ex(points, k, num_indexed_dims=nd, beta=0.025)
    feats, wf = _lib.radius_features(points, k, nd=nd)
    dk = dists[:, k].astype(np.float64)
    rng = np.random.default_rng(seed)
    sel = rng.choice(len(points), min(max_queries, len(points)), replace=False)
    ok = dk[sel] > 0
    sel = sel[ok]
    y = np.log(dk[sel] / wf).astype(np.float32)
    return feats[sel], y


def default_sets(nd: int, n: int, seed: int, d: int | None = None):
    """Seeded training point sets of `d` dims (default nd): for 2-D, the uniform box and
    three uniform/clustered mixes; above, the uniform and the skewed (exponential) cube."""
    d = nd if d is None else d
    out = []
    if d == 2:
        out.append(synthetic.uniform_box(n, seed=seed + 1))
        out.append(synthetic.mixed_box(n, seed=seed + 2))
        out.append(synthetic.mixed_box(n, seed=seed + 3, frac_uniform=0.5, nclusters=16, sigma=1.0))
        out.append(synthetic.mixed_box(n, seed=seed + 4, frac_uniform=0.9, nclusters=128, sigma=0.25))
    else:
        out.append(synthetic.unit_cube(n, d, seed=seed + 1))
        out.append(synthetic.skewed_cube(n, d, seed=seed + 2))
    return out


def build_dataset(nd: int, n: int, ks, per_set: int, seed: int, log=print, d: int | None = None):
    X, Y = [], []
    for si, pts in enumerate(default_sets(nd, n, seed, d)):
        for k in ks:
            t0 = time.time()
            f, y = training_pairs(pts, k, nd, per_set, seed + 100 * si + k)
            X.append(f)
            Y.append(y)
            log(f"[train] set {si} k={k}: {len(y)} pairs ({time.time() - t0:.1f} s)")
    return np.concatenate(X), np.concatenate(Y)


def calibrate_margin(w: np.ndarray, X: np.ndarray, Y: np.ndarray, quantile: float) -> float:
    pred = _lib.mlp_forward(w, X)
    return float(np.quantile(Y - pred, quantile))


def train(nd: int = 2, n: int = 2_000_000, ks=(8, 16, 32, 64, 128), per_set: int = 200_000,
          epochs: int = 30, batch: int = 4096, lr: float = 2e-3, seed: int = 20261202,
          quantile: float = 0.97, out_dir: str = WEIGHTS_DIR, log=print, d: int | None = None) -> dict:
    d = nd if d is None else d
    X, Y = build_dataset(nd, n, ks, per_set, seed, log, d)
    perm = np.random.default_rng(seed).permutation(len(Y))
    nval = max(1, len(Y) // 10)
    val, tr = perm[:nval], perm[nval:]
    w0 = init_weights(seed)
    t0 = time.time()
    w, losses = _lib.mlp_train(w0, X[tr], Y[tr], epochs, batch, lr, seed)
    t_train = time.time() - t0
    pred_val = _lib.mlp_forward(w, X[val])
    mse_val = float(np.mean((pred_val - Y[val]) ** 2))
    base_val = float(np.mean((Y[val] - Y[tr].mean()) ** 2))
    margin = calibrate_margin(w, X[tr], Y[tr], quantile)
    cover = float(np.mean(Y[val] - pred_val <= margin))
    info = {"d": d, "nd": nd, "pairs": int(len(Y)), "epochs": epochs, "batch": batch, "lr": lr,
            "seed": seed, "quantile": quantile, "margin": margin, "val_mse": mse_val,
            "val_mse_constant_predictor": base_val, "val_coverage": cover,
            "train_seconds": t_train, "loss_per_epoch": [float(v) for v in losses],
            "training_sets": f"{len(default_sets(nd, 16, seed, d))} seeded {d}-D sets of {n} points, k in {list(ks)}"}
    save(nd, w, margin, info, out_dir, d)
    log(f"[train] d={d} nd={nd}: val mse {mse_val:.4f} (constant predictor {base_val:.4f}), "
        f"margin {margin:.3f} covers {cover:.3f} of validation queries; {t_train:.1f} s")
    return info


def load(nd: int, directory: str = WEIGHTS_DIR, d: int | None = None):
    path = weights_path(nd, directory, d)
    if not os.path.exists(path):
        return None
    z = np.load(path, allow_pickle=False)
    return z["weights"].astype(np.float32), float(z["margin"]), json.loads(str(z["info"]))


def register_default_models(directory: str = WEIGHTS_DIR) -> list:
    """Registers every weights file in `directory`; returns the (d, nd) pairs loaded (d = 0:
    the model applies to any NDIM with nd indexed dims)."""
    done = []
    if not os.path.isdir(directory):
        return done
    for name in sorted(os.listdir(directory)):
        m = re.fullmatch(r"radius_mlp_(?:d(\d+)_)?nd(\d+)\.npz", name)
        if not m:
            continue
        nd = int(m.group(2))
        d = int(m.group(1)) if m.group(1) else 0
        w = load(nd, directory, d if d else None)
        if w is None:
            continue
        _lib.set_radius_model(nd, w[0], w[1], ndim=d)
        done.append((d, nd))
    return done


def has_model(done: list, d: int, nd: int) -> bool:
    return (d, nd) in done or (0, nd) in done


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--nd", type=int, default=2)
    ap.add_argument("--d", type=int, default=None, help="NDIM of the training data (default nd)")
    ap.add_argument("--n", type=int, default=2_000_000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--per-set", type=int, default=200_000)
    ap.add_argument("--out", default=WEIGHTS_DIR)
    args = ap.parse_args()
    d = args.nd if args.d is None else args.d
    ks = (8, 16, 32, 64, 128) if d == 2 else (16, 32, 64)
    train(nd=args.nd, n=args.n, ks=ks, per_set=args.per_set, epochs=args.epochs, out_dir=args.out, d=d)


if __name__ == "__main__":
    main()
