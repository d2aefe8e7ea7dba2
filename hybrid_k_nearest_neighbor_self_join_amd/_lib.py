"""ctypes binding of libgpuknnselfjoin.so (the C ABI declared in include/gpuknnselfjoin.h).

The shared library is built in-tree by `build.py`; importing this module never falls back to
another implementation: a missing library is built once with hipcc, and a library that cannot
be loaded raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgpuknnselfjoin.so")
# the DTYPE=double build: its KNNJoinPy takes double* (params.h:5)
LIB_DOUBLE_PATH = os.path.join(PKG, "double", "libgpuknnselfjoin.so")

KNN_OK, KNN_ERR_ARG, KNN_ERR_TOO_FEW, KNN_ERR_EPS_ZERO, KNN_ERR_NO_PROGRESS = 0, 1, 2, 3, 4
KNN_ERR_HIP, KNN_ERR_NOMEM, KNN_ERR_NO_DEVICE, KNN_ERR_MODEL = 5, 6, 7, 8
KNN_FLOAT32, KNN_FLOAT64 = 0, 1
KNN_MODE_BETA, KNN_MODE_LEARNED, KNN_MODE_BRUTE = 0, 1, 2

STATUS_NAMES = {
    0: "KNN_OK", 1: "KNN_ERR_ARG", 2: "KNN_ERR_TOO_FEW", 3: "KNN_ERR_EPS_ZERO",
    4: "KNN_ERR_NO_PROGRESS", 5: "KNN_ERR_HIP", 6: "KNN_ERR_NOMEM", 7: "KNN_ERR_NO_DEVICE",
    8: "KNN_ERR_MODEL",
}

# every symbol include/gpuknnselfjoin.h declares
EXPORTED_SYMBOLS = [
    "KNNJoinPy", "KNNJoinPyDouble", "knnjoin_py_dtype", "knnjoin_ex", "knnjoin_ex_noalloc", "knnjoin_default_params",
    "knnjoin_last_status", "knnjoin_last_error", "knnjoin_release", "knnjoin_stage_reorder",
    "knnjoin_stage_estimate_eps", "knnjoin_stage_index", "knnjoin_stage_join_pass",
    "knnjoin_set_radius_model", "knnjoin_set_radius_model_dims", "knnjoin_radius_features",
    "knnjoin_mlp_forward",
    "knnjoin_mlp_grad", "knnjoin_mlp_train", "knnjoin_batch_default_params", "knnjoin_batched",
    "knnjoin_batch_autotune", "knnjoin_outlier_scores",
]
MLP_NW = 5057


class KnnParams(ctypes.Structure):
    _fields_ = [
        ("num_indexed_dims", ctypes.c_uint32),
        ("beta", ctypes.c_double),
        ("mode", ctypes.c_int32),
        ("reorder", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("device_ptrs", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
        ("brute_fallback", ctypes.c_int32),
        ("hybrid_threads", ctypes.c_uint32),
        ("hybrid_max_queries", ctypes.c_uint32),
    ]


class KnnStats(ctypes.Structure):
    _fields_ = [
        ("perm", ctypes.c_uint32 * 16),
        ("eps0", ctypes.c_float),
        ("eps_final", ctypes.c_float),
        ("iterations", ctypes.c_uint32),
        ("reindexes", ctypes.c_uint32),
        ("kernel_launches", ctypes.c_uint32),
        ("failed_final", ctypes.c_uint32),
        ("candidates", ctypes.c_uint64),
        ("algo_bytes", ctypes.c_uint64),
        ("checksum", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("ms_reorder", ctypes.c_double),
        ("ms_eps", ctypes.c_double),
        ("ms_index", ctypes.c_double),
        ("ms_join", ctypes.c_double),
        ("ms_other", ctypes.c_double),
        ("ms_range_kernel", ctypes.c_double),
        ("ms_range_kernel_max", ctypes.c_double),
        ("algo_bytes_max", ctypes.c_uint64),
        ("overflow_queries", ctypes.c_uint64),
        ("eps_bucket", ctypes.c_uint32),
        ("brute_queries", ctypes.c_uint32),
        ("ms_index_first", ctypes.c_double),
        ("ms_reference_total", ctypes.c_double),
        ("host_syncs", ctypes.c_uint32),
        ("hybrid_queries", ctypes.c_uint32),
        ("ms_hybrid_host", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "perm" else v
        return d


class KnnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def lib() -> ctypes.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                from . import build as _build
                _build.build()
            L = ctypes.CDLL(LIB_PATH)
            _setup(L)
            _lib = L
    return _lib


def _setup(L):
    vp, u32, u64, i32, dbl, flt = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_float)
    L.KNNJoinPy.argtypes = [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, vp, vp]
    L.KNNJoinPy.restype = None
    L.KNNJoinPyDouble.argtypes = [vp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, vp, vp]
    L.KNNJoinPyDouble.restype = None
    L.knnjoin_py_dtype.argtypes = []
    L.knnjoin_py_dtype.restype = i32
    L.knnjoin_ex.argtypes = [vp, i32, u32, u32, u32, ctypes.POINTER(KnnParams), vp, vp,
                             ctypes.POINTER(KnnStats)]
    L.knnjoin_ex.restype = i32
    L.knnjoin_ex_noalloc.argtypes = [vp, i32, u32, u32, u32, ctypes.POINTER(KnnParams), vp, vp,
                                     ctypes.POINTER(KnnStats)]
    L.knnjoin_ex_noalloc.restype = i32
    L.knnjoin_default_params.argtypes = [ctypes.POINTER(KnnParams)]
    L.knnjoin_default_params.restype = None
    L.knnjoin_last_status.argtypes = []
    L.knnjoin_last_status.restype = i32
    L.knnjoin_last_error.argtypes = []
    L.knnjoin_last_error.restype = ctypes.c_char_p
    L.knnjoin_release.argtypes = []
    L.knnjoin_release.restype = None
    L.knnjoin_stage_reorder.argtypes = [vp, u32, u32, vp, vp]
    L.knnjoin_stage_reorder.restype = i32
    L.knnjoin_stage_estimate_eps.argtypes = [vp, u32, u32, u32, dbl, vp, vp, vp, vp, vp]
    L.knnjoin_stage_estimate_eps.restype = i32
    L.knnjoin_stage_index.argtypes = [vp, u32, u32, u32, flt, vp, vp, vp, vp, vp, vp, vp]
    L.knnjoin_stage_index.restype = i32
    L.knnjoin_stage_join_pass.argtypes = [vp, u32, u32, u32, u32, flt, vp, u32, vp, vp, vp, vp]
    L.knnjoin_stage_join_pass.restype = i32
    L.knnjoin_set_radius_model.argtypes = [u32, vp, u32, flt]
    L.knnjoin_set_radius_model.restype = i32
    L.knnjoin_set_radius_model_dims.argtypes = [u32, u32, vp, u32, flt]
    L.knnjoin_set_radius_model_dims.restype = i32
    L.knnjoin_radius_features.argtypes = [vp, u32, u32, u32, u32, i32, vp, vp, vp]
    L.knnjoin_radius_features.restype = i32
    L.knnjoin_mlp_forward.argtypes = [vp, vp, u32, vp]
    L.knnjoin_mlp_forward.restype = i32
    L.knnjoin_mlp_grad.argtypes = [vp, vp, vp, u32, vp, vp]
    L.knnjoin_mlp_grad.restype = i32
    L.knnjoin_mlp_train.argtypes = [vp, vp, vp, u32, u32, u32, flt, u64, vp]
    L.knnjoin_mlp_train.restype = i32
    L.knnjoin_outlier_scores.argtypes = [vp, vp, u32, u32, i32, vp, vp, vp, vp]
    L.knnjoin_outlier_scores.restype = i32


def p(a) -> ctypes.c_void_p | None:
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return a.ctypes.data_as(ctypes.c_void_p)


def check(rc: int):
    if rc != KNN_OK:
        msg = lib().knnjoin_last_error()
        raise KnnError(rc, msg.decode() if msg else "")


def default_params(**kw) -> KnnParams:
    prm = KnnParams()
    lib().knnjoin_default_params(ctypes.byref(prm))
    for k, v in kw.items():
        setattr(prm, k, v)
    return prm


# ------------------------------------------------------------------ convenience wrappers
def knnjoin_ex(points: np.ndarray, k: int, *, num_indexed_dims: int = 0, beta: float = 0.05,
               mode: int = KNN_MODE_BETA, reorder: int = 1, verbose: int = 0,
               brute_fallback: int = -1, hybrid_threads: int = 0, hybrid_max_queries: int = 0):
    """Whole join through knnjoin_ex on host arrays.  Returns (ids[N,K+1], dists[N,K+1], stats)."""
    pts = np.ascontiguousarray(points)
    if pts.dtype not in (np.float32, np.float64):
        pts = pts.astype(np.float32)
    n, d = pts.shape
    dt = KNN_FLOAT32 if pts.dtype == np.float32 else KNN_FLOAT64
    ids = np.full((n, k + 1), -1, np.int32)
    dists = np.zeros((n, k + 1), pts.dtype)
    prm = default_params(num_indexed_dims=num_indexed_dims, beta=beta, mode=mode,
                         reorder=reorder, verbose=verbose, brute_fallback=brute_fallback,
                         hybrid_threads=hybrid_threads, hybrid_max_queries=hybrid_max_queries)
    st = KnnStats()
    rc = lib().knnjoin_ex(p(pts), dt, n, d, k, ctypes.byref(prm), p(ids), p(dists),
                          ctypes.byref(st))
    check(rc)
    return ids, dists, st


def stage_reorder(points: np.ndarray):
    pts = np.ascontiguousarray(points, np.float32)
    n, d = pts.shape
    perm = np.zeros(d, np.uint32)
    var = np.zeros(d, np.float32)
    check(lib().knnjoin_stage_reorder(p(pts), n, d, p(perm), p(var)))
    return perm, var


def stage_estimate_eps(points: np.ndarray, k: int, beta: float):
    pts = np.ascontiguousarray(points, np.float32)
    n, d = pts.shape
    eps = np.zeros(1, np.float32)
    avg = np.zeros(1, np.float64)
    cnt = np.zeros(1, np.uint64)
    bucket = np.zeros(1, np.uint32)
    hist = np.zeros(10000, np.uint32)
    check(lib().knnjoin_stage_estimate_eps(p(pts), n, d, k, beta, p(eps), p(avg), p(cnt),
                                           p(bucket), p(hist)))
    return float(eps[0]), float(avg[0]), int(cnt[0]), int(bucket[0]), hist


def stage_index(points: np.ndarray, nd: int, eps: float):
    pts = np.ascontiguousarray(points, np.float32)
    n, d = pts.shape
    minArr = np.zeros(nd, np.float32)
    nCells = np.zeros(nd, np.uint32)
    tot = np.zeros(1, np.uint64)
    sortedIds = np.zeros(n, np.uint32)
    cellKeys = np.zeros(n, np.uint64)
    cellStart = np.zeros(n + 1, np.uint32)
    ncell = np.zeros(1, np.uint32)
    check(lib().knnjoin_stage_index(p(pts), n, d, nd, eps, p(minArr), p(nCells), p(tot),
                                    p(sortedIds), p(cellKeys), p(cellStart), p(ncell)))
    c = int(ncell[0])
    return minArr, nCells, int(tot[0]), sortedIds, cellKeys[:c].copy(), cellStart[:c + 1].copy()


def stage_join_pass(points: np.ndarray, nd: int, k: int, eps: float, queries: np.ndarray):
    pts = np.ascontiguousarray(points, np.float32)
    n, d = pts.shape
    q = np.ascontiguousarray(queries, np.uint32)
    ids = np.full((n, k + 1), -1, np.int32)
    dists = np.zeros((n, k + 1), np.float32)
    failed = np.zeros(n, np.uint8)
    nf = np.zeros(1, np.uint32)
    check(lib().knnjoin_stage_join_pass(p(pts), n, d, nd, k, eps, p(q), len(q), p(ids), p(dists),
                                        p(failed), p(nf)))
    return ids, dists, failed, int(nf[0])


# ------------------------------------------------------------------ learned radius
def set_radius_model(nd: int, weights: np.ndarray, margin: float, ndim: int = 0):
    """Registers a radius model for `nd` indexed dims: for any NDIM (ndim 0) or for NDIM =
    ndim only (preferred over the any-NDIM model)."""
    w = np.ascontiguousarray(weights, np.float32).reshape(-1)
    check(lib().knnjoin_set_radius_model_dims(ndim, nd, p(w), len(w), margin))


def radius_features(points: np.ndarray, k: int, nd: int | None = None, reorder: int = 1,
                    want_radius: bool = False):
    pts = np.ascontiguousarray(points, np.float32)
    n, d = pts.shape
    nd = min(d, 6) if nd is None else nd
    feats = np.zeros((n, 12), np.float32)
    wf = np.zeros(1, np.float32)
    rad = np.zeros(n, np.float32) if want_radius else None
    check(lib().knnjoin_radius_features(p(pts), n, d, k, nd, reorder, p(feats), p(wf), p(rad)))
    return (feats, float(wf[0]), rad) if want_radius else (feats, float(wf[0]))


def mlp_forward(weights: np.ndarray, X: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(weights, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros(len(X), np.float32)
    check(lib().knnjoin_mlp_forward(p(w), p(X), len(X), p(out)))
    return out


def mlp_grad(weights: np.ndarray, X: np.ndarray, Y: np.ndarray):
    w = np.ascontiguousarray(weights, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    g = np.zeros(MLP_NW, np.float32)
    loss = np.zeros(1, np.float32)
    check(lib().knnjoin_mlp_grad(p(w), p(X), p(Y), len(X), p(g), p(loss)))
    return g, float(loss[0])


def mlp_train(weights: np.ndarray, X: np.ndarray, Y: np.ndarray, epochs: int, batch: int,
              lr: float, seed: int):
    w = np.array(weights, np.float32, copy=True)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    losses = np.zeros(epochs, np.float32)
    check(lib().knnjoin_mlp_train(p(w), p(X), p(Y), len(X), epochs, batch, lr, seed, p(losses)))
    return w, losses


def outlier_scores(ids: np.ndarray, dists: np.ndarray):
    """Outlier scores of an N x (K+1) table (host arrays): (dist_sum float32, dist_rank,
    indegree, indegree_rank uint32); see knnjoin_outlier_scores."""
    ids = np.ascontiguousarray(ids, np.int32)
    dists = np.ascontiguousarray(dists, np.float32)
    n, k1 = ids.shape
    ds = np.zeros(n, np.float32)
    dr, dg, gr = (np.zeros(n, np.uint32) for _ in range(3))
    check(lib().knnjoin_outlier_scores(p(ids), p(dists), n, k1, 0, p(ds), p(dr), p(dg), p(gr)))
    return ds, dr, dg, gr
