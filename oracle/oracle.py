"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the reported CPU baseline.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

ORC_OK, ORC_ERR_ARG, ORC_ERR_TOO_FEW, ORC_ERR_EPS_ZERO, ORC_ERR_NO_PROGRESS = range(5)
NBUCKETS = 10000


class EpsInfo(ctypes.Structure):
    _fields_ = [
        ("avg", ctypes.c_double),
        ("cnt", ctypes.c_uint64),
        ("bucket_width", ctypes.c_double),
        ("offset", ctypes.c_uint32),
        ("nsample", ctypes.c_uint32),
        ("bucket", ctypes.c_uint32),
        ("threshold", ctypes.c_uint32),
        ("eps", ctypes.c_float),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("perm", ctypes.c_uint32 * 16),
        ("eps0", ctypes.c_float),
        ("eps_final", ctypes.c_float),
        ("iterations", ctypes.c_uint32),
        ("reindexes", ctypes.c_uint32),
        ("candidates", ctypes.c_uint64),
        ("checksum", ctypes.c_double),
        ("ms_prep", ctypes.c_double),
        ("ms_index", ctypes.c_double),
        ("ms_passes", ctypes.c_double),
    ]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _setup(L):
    vp, u32, u64, i32, dbl, flt = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_float)
    L.orc_reorder_f32.argtypes = [vp, u32, u32, vp, vp]
    L.orc_apply_perm_f32.argtypes = [vp, u32, u32, vp, vp]
    L.orc_avg_dist_f32.argtypes = [vp, u32, u32, vp, vp]
    L.orc_estimate_eps_f32.argtypes = [vp, u32, u32, u32, dbl, vp, ctypes.POINTER(EpsInfo)]
    L.orc_estimate_eps_f32.restype = i32
    L.orc_minmax_f32.argtypes = [vp, u32, u32, u32, vp, vp]
    L.orc_grid_dims_f32.argtypes = [vp, vp, u32, flt, vp, vp, vp]
    L.orc_build_index_f32.argtypes = [vp, u32, u32, u32, flt, vp, vp, vp, vp, vp, vp]
    L.orc_build_index_f32.restype = u32
    L.orc_join_pass_f32.argtypes = [vp, u32, u32, u32, u32, flt, vp, vp, u32, vp, vp, vp, vp,
                                    u32, i32, vp, vp, vp, vp]
    L.orc_join_pass_f32.restype = u32
    L.orc_knnjoin_f32.argtypes = [vp, u32, u32, u32, u32, dbl, i32, vp, vp, ctypes.POINTER(Stats)]
    L.orc_knnjoin_f32.restype = i32
    L.orc_knnjoin_subset_f32.argtypes = [vp, u32, u32, u32, u32, dbl, i32, vp, u32, vp, vp,
                                         ctypes.POINTER(Stats)]
    L.orc_knnjoin_subset_f32.restype = i32
    L.orc_bruteforce_f32.argtypes = [vp, u32, u32, u32, vp, u32, i32, vp, vp]
    L.orc_checksum_f32.argtypes = [vp, u64, u32]
    L.orc_checksum_f32.restype = dbl
    L.orc_outlier_scores.argtypes = [vp, vp, u32, u32, vp, vp, vp, vp]
    L.orc_outlier_scores.restype = None
    L.orc_reorder_f64.argtypes = [vp, u32, u32, vp, vp]
    L.orc_estimate_eps_f64.argtypes = [vp, u32, u32, u32, dbl, vp, ctypes.POINTER(EpsInfo)]
    L.orc_estimate_eps_f64.restype = i32
    L.orc_knnjoin_f64.argtypes = [vp, u32, u32, u32, u32, dbl, i32, vp, vp, ctypes.POINTER(Stats)]
    L.orc_knnjoin_f64.restype = i32
    L.orc_bruteforce_f64.argtypes = [vp, u32, u32, u32, vp, u32, i32, vp, vp]
    L.orc_grow_eps_f32.argtypes = [flt, flt]
    L.orc_grow_eps_f32.restype = flt
    L.orc_cell_coord_f32.argtypes = [flt, flt, flt, u32]
    L.orc_cell_coord_f32.restype = u32
    L.orc_mlp_forward.argtypes = [vp, vp, u32, vp]
    L.orc_mlp_grad.argtypes = [vp, vp, vp, u32, vp, vp]
    L.orc_mlp_train.argtypes = [vp, vp, vp, u32, u32, u32, flt, u64]
    L.orc_radius_features.argtypes = [vp, u32, u32, u32, u32, vp, vp]


def _threads(nthreads):
    return int(nthreads or os.cpu_count() or 1)


# ---------------------------------------------------------------- stage functions
def reorder(pts: np.ndarray):
    pts = np.ascontiguousarray(pts)
    n, d = pts.shape
    perm = np.zeros(d, np.uint32)
    var = np.zeros(d, pts.dtype)
    f = lib().orc_reorder_f32 if pts.dtype == np.float32 else lib().orc_reorder_f64
    f(_p(pts), n, d, _p(perm), _p(var))
    return perm, var


def apply_perm(pts: np.ndarray, perm: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(pts[:, perm])


def avg_dist(pts: np.ndarray):
    pts = np.ascontiguousarray(pts, np.float32)
    n, d = pts.shape
    avg = np.zeros(1, np.float64)
    cnt = np.zeros(1, np.uint64)
    lib().orc_avg_dist_f32(_p(pts), n, d, _p(avg), _p(cnt))
    return float(avg[0]), int(cnt[0])


def estimate_eps(pts: np.ndarray, k: int, beta: float):
    pts = np.ascontiguousarray(pts)
    n, d = pts.shape
    hist = np.zeros(NBUCKETS, np.uint32)
    info = EpsInfo()
    f = lib().orc_estimate_eps_f32 if pts.dtype == np.float32 else lib().orc_estimate_eps_f64
    rc = f(_p(pts), n, d, k, beta, _p(hist), ctypes.byref(info))
    return rc, info, hist


def minmax(pts: np.ndarray, nd: int):
    pts = np.ascontiguousarray(pts, np.float32)
    n, d = pts.shape
    mn = np.zeros(nd, np.float32)
    mx = np.zeros(nd, np.float32)
    lib().orc_minmax_f32(_p(pts), n, d, nd, _p(mn), _p(mx))
    return mn, mx


def grid_dims(mn, mx, eps: float):
    nd = len(mn)
    mn = np.ascontiguousarray(mn, np.float32)
    mx = np.ascontiguousarray(mx, np.float32)
    minArr = np.zeros(nd, np.float32)
    nCells = np.zeros(nd, np.uint32)
    tot = np.zeros(1, np.uint64)
    lib().orc_grid_dims_f32(_p(mn), _p(mx), nd, eps, _p(minArr), _p(nCells), _p(tot))
    return minArr, nCells, int(tot[0])


def build_index(pts: np.ndarray, nd: int, eps: float, minArr, nCells):
    pts = np.ascontiguousarray(pts, np.float32)
    n, d = pts.shape
    pointKey = np.zeros(n, np.uint64)
    sortedIds = np.zeros(n, np.uint32)
    cellKeys = np.zeros(n, np.uint64)
    cellStart = np.zeros(n + 1, np.uint32)
    c = lib().orc_build_index_f32(_p(pts), n, d, nd, eps, _p(np.ascontiguousarray(minArr, np.float32)),
                                  _p(np.ascontiguousarray(nCells, np.uint32)), _p(pointKey),
                                  _p(sortedIds), _p(cellKeys), _p(cellStart))
    return pointKey, sortedIds, cellKeys[:c].copy(), cellStart[:c + 1].copy()


def grow_eps(eps: float, eps0: float) -> float:
    return float(lib().orc_grow_eps_f32(eps, eps0))


def cell_coord(x: float, mn: float, eps: float, ncell: int) -> int:
    return int(lib().orc_cell_coord_f32(x, mn, eps, ncell))


# ---------------------------------------------------------------- whole join / brute force
def knnjoin(pts: np.ndarray, k: int, nd: int | None = None, beta: float = 0.05,
            nthreads: int | None = None, subset: np.ndarray | None = None):
    """Plain method (reorder, BETA epsilon, grid, failed-query loop).  Returns
    (rc, ids, dists, Stats): ids/dists are [N, K+1] for the whole join; with `subset` only
    the subset's queries are joined and ids/dists are [len(subset), K+1], row i = point
    subset[i]."""
    pts = np.ascontiguousarray(pts)
    n, d = pts.shape
    nd = d if nd is None else nd
    st = Stats()
    if subset is not None:
        if pts.dtype != np.float32:
            raise ValueError("subset joins take float32 points")
        sub = np.ascontiguousarray(subset, np.uint32)
        # rows are written only for the subset: untouched pages of np.empty are never mapped
        ids = np.empty((n, k + 1), np.int32)
        dists = np.empty((n, k + 1), np.float32)
        rc = lib().orc_knnjoin_subset_f32(_p(pts), n, d, k, nd, beta, _threads(nthreads),
                                          _p(sub), len(sub), _p(ids), _p(dists), ctypes.byref(st))
        return rc, ids[sub].copy(), dists[sub].copy(), st
    ids = np.full((n, k + 1), -1, np.int32)
    dists = np.zeros((n, k + 1), pts.dtype)
    if pts.dtype == np.float32:
        rc = lib().orc_knnjoin_f32(_p(pts), n, d, k, nd, beta, _threads(nthreads), _p(ids),
                                   _p(dists), ctypes.byref(st))
    else:
        rc = lib().orc_knnjoin_f64(_p(pts), n, d, k, nd, beta, _threads(nthreads), _p(ids),
                                   _p(dists), ctypes.byref(st))
    return rc, ids, dists, st


def bruteforce(pts: np.ndarray, k: int, queries: np.ndarray | None = None,
               nthreads: int | None = None):
    """FP64 brute force over every point; rows follow `queries` (default: all points).
    Distances follow `pts`' column order, so pass the reordered array to compare
    distances bitwise with the join."""
    pts = np.ascontiguousarray(pts)
    n, d = pts.shape
    q = np.arange(n, dtype=np.uint32) if queries is None else np.ascontiguousarray(queries, np.uint32)
    ids = np.zeros((len(q), k + 1), np.int32)
    dists = np.zeros((len(q), k + 1), pts.dtype)
    f = lib().orc_bruteforce_f32 if pts.dtype == np.float32 else lib().orc_bruteforce_f64
    f(_p(pts), n, d, k, _p(q), len(q), _threads(nthreads), _p(ids), _p(dists))
    return ids, dists


def checksum(dists: np.ndarray) -> float:
    dists = np.ascontiguousarray(dists, np.float32)
    return float(lib().orc_checksum_f32(_p(dists), dists.shape[0], dists.shape[1]))


def outlier_scores(ids: np.ndarray, dists: np.ndarray):
    """(dist_sum, dist_rank, indegree, indegree_rank) of an N x (K+1) table."""
    ids = np.ascontiguousarray(ids, np.int32)
    dists = np.ascontiguousarray(dists, np.float32)
    n, k1 = ids.shape
    ds = np.zeros(n, np.float32)
    dr, dg, gr = (np.zeros(n, np.uint32) for _ in range(3))
    lib().orc_outlier_scores(_p(ids), _p(dists), n, k1, _p(ds), _p(dr), _p(dg), _p(gr))
    return ds, dr, dg, gr


def join_pass(pts_r: np.ndarray, nd: int, k: int, eps: float, queries: np.ndarray,
              nthreads: int | None = None):
    """One range pass at a fixed eps over `queries` (points already reordered), on the
    index built at eps from the data's min/max.  Returns (ids, dists, failed, nfail)."""
    pts = np.ascontiguousarray(pts_r, np.float32)
    n, d = pts.shape
    mn, mx = minmax(pts, nd)
    minArr, nCells, _ = grid_dims(mn, mx, eps)
    _, sortedIds, cellKeys, cellStart = build_index(pts, nd, eps, minArr, nCells)
    q = np.ascontiguousarray(queries, np.uint32)
    ids = np.full((n, k + 1), -1, np.int32)
    dists = np.zeros((n, k + 1), np.float32)
    failed = np.zeros(n, np.uint8)
    cand = np.zeros(1, np.uint64)
    nf = lib().orc_join_pass_f32(_p(pts), n, d, nd, k, eps, _p(minArr), _p(nCells), len(cellKeys),
                                 _p(cellKeys), _p(cellStart), _p(sortedIds), _p(q), len(q),
                                 _threads(nthreads), _p(ids), _p(dists), _p(failed), _p(cand))
    return ids, dists, failed, int(nf)


# ---------------------------------------------------------------- learned radius
MLP_NW = 5057


def mlp_forward(W, X):
    W = np.ascontiguousarray(W, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros(len(X), np.float32)
    lib().orc_mlp_forward(_p(W), _p(X), len(X), _p(out))
    return out


def mlp_grad(W, X, Y):
    W = np.ascontiguousarray(W, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    G = np.zeros(MLP_NW, np.float32)
    loss = np.zeros(1, np.float32)
    lib().orc_mlp_grad(_p(W), _p(X), _p(Y), len(X), _p(G), _p(loss))
    return G, float(loss[0])


def mlp_train(W, X, Y, epochs, batch, lr, seed):
    W = np.array(W, np.float32, copy=True)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    lib().orc_mlp_train(_p(W), _p(X), _p(Y), len(X), epochs, batch, lr, seed)
    return W


def radius_features(pts_r, k, nd):
    pts = np.ascontiguousarray(pts_r, np.float32)
    n, d = pts.shape
    feat = np.zeros((n, 12), np.float32)
    wf = np.zeros(1, np.float32)
    lib().orc_radius_features(_p(pts), n, d, k, nd, _p(feat), _p(wf))
    return feat, float(wf[0])
