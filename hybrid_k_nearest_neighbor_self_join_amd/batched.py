"""Batched joins of many independent point sets (knnjoin_batched / knnjoin_batch_autotune).

Each set gets exactly the rows `knnjoin_ex` gives it alone; the library joins all sets of one
dimensionality together (one index, one persistent range-kernel launch per radius pass) and
replays a captured hipGraph of the whole join for calls with the same set shapes.

    ids, dists, stats = batched.knnjoin_batched([pts0, pts1, ...], [k0, k1, ...])
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib

TUNE_CACHE = os.path.join(_lib.PKG, "batch_tuning.json")


class BatchParams(ctypes.Structure):
    _fields_ = [
        ("num_indexed_dims", ctypes.c_uint32),
        ("reorder", ctypes.c_int32),
        ("device_ptrs", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("graph_passes", ctypes.c_uint32),
        ("tile_queries", ctypes.c_uint32),
        ("grid_blocks", ctypes.c_uint32),
        ("verbose", ctypes.c_int32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("ms_device", ctypes.c_double),
        ("ms_setup", ctypes.c_double),
        ("groups", ctypes.c_uint32),
        ("graph_replayed", ctypes.c_uint32),
        ("passes_graph", ctypes.c_uint32),
        ("passes_eager", ctypes.c_uint32),
        ("tile_queries", ctypes.c_uint32),
        ("grid_blocks", ctypes.c_uint32),
        ("points", ctypes.c_uint64),
        ("candidates", ctypes.c_uint64),
        ("algo_bytes", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_bound = False


def _L():
    global _bound
    L = _lib.lib()
    if not _bound:
        vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
        L.knnjoin_batch_default_params.argtypes = [ctypes.POINTER(BatchParams)]
        L.knnjoin_batch_default_params.restype = None
        L.knnjoin_batched.argtypes = [u32, vp, vp, vp, vp, ctypes.POINTER(BatchParams), vp, vp,
                                      ctypes.POINTER(BatchStats)]
        L.knnjoin_batched.restype = i32
        L.knnjoin_batch_autotune.argtypes = [u32, vp, vp, vp, vp, ctypes.POINTER(BatchParams), vp, vp,
                                             vp, vp, vp, vp]
        L.knnjoin_batch_autotune.restype = i32
        _bound = True
    return L


def default_params(**kw) -> BatchParams:
    prm = BatchParams()
    _L().knnjoin_batch_default_params(ctypes.byref(prm))
    for k, v in kw.items():
        setattr(prm, k, v)
    return prm


class _Args:
    """ctypes arrays of the per-set pointers and sizes (host numpy arrays or device pointers)."""

    def __init__(self, sets, ks, outs_ids=None, outs_dists=None, device=False):
        m = len(sets)
        self.m = m
        if device:
            ptrs = [int(s[0]) for s in sets]
            self.n = np.array([s[1] for s in sets], np.uint32)
            self.d = np.array([s[2] for s in sets], np.uint32)
            self.keep = []
        else:
            arrs = [np.ascontiguousarray(s, np.float32) for s in sets]
            self.keep = arrs
            ptrs = [a.ctypes.data for a in arrs]
            self.n = np.array([a.shape[0] for a in arrs], np.uint32)
            self.d = np.array([a.shape[1] for a in arrs], np.uint32)
        self.k = np.array(ks, np.uint32)
        self.pts = (ctypes.c_void_p * m)(*ptrs)
        if outs_ids is None:
            self.ids = [np.full((int(n), int(k) + 1), -1, np.int32) for n, k in zip(self.n, self.k)]
            self.dists = [np.zeros((int(n), int(k) + 1), np.float32) for n, k in zip(self.n, self.k)]
            ip = [a.ctypes.data for a in self.ids]
            dp = [a.ctypes.data for a in self.dists]
        else:
            self.ids, self.dists = outs_ids, outs_dists
            ip = [int(x) if isinstance(x, int) else x.ctypes.data for x in outs_ids]
            dp = [int(x) if isinstance(x, int) else x.ctypes.data for x in outs_dists]
        self.pids = (ctypes.c_void_p * m)(*ip)
        self.pdists = (ctypes.c_void_p * m)(*dp)

    def call_args(self):
        return (self.m, ctypes.cast(self.pts, ctypes.c_void_p), _lib.p(self.n), _lib.p(self.d),
                _lib.p(self.k))


def knnjoin_batched(sets, ks, *, num_indexed_dims: int = 0, use_graph: bool = True,
                    graph_passes: int = 0, tile_queries: int = 0, grid_blocks: int = 0,
                    reorder: int = 1, verbose: int = 0, device=False, out_ids=None, out_dists=None):
    """Joins every set.  `sets`: host float32 arrays (n_i x d_i), or with device=True tuples
    (device pointer, n_i, d_i) plus out_ids/out_dists lists of device pointers.
    Returns (ids list, dists list, BatchStats)."""
    a = _Args(sets, ks, out_ids, out_dists, device)
    prm = default_params(num_indexed_dims=num_indexed_dims, use_graph=int(use_graph),
                         graph_passes=graph_passes, tile_queries=tile_queries,
                         grid_blocks=grid_blocks, reorder=reorder, verbose=verbose,
                         device_ptrs=int(device))
    st = BatchStats()
    rc = _L().knnjoin_batched(*a.call_args(), ctypes.byref(prm), ctypes.cast(a.pids, ctypes.c_void_p),
                              ctypes.cast(a.pdists, ctypes.c_void_p), ctypes.byref(st))
    _lib.check(rc)
    return a.ids, a.dists, st


def autotune(sets, ks, *, num_indexed_dims: int = 0, device=False, out_ids=None, out_dists=None,
             cache: str | None = TUNE_CACHE):
    """Times the tile-size x grid-size candidates on these sets and keeps the fastest in the
    library for later calls with the same shapes.  Returns (tile_queries, grid_blocks,
    {"tq/grid": ms}).  With `cache`, the choice is also stored under a key of the set shapes."""
    a = _Args(sets, ks, out_ids, out_dists, device)
    prm = default_params(num_indexed_dims=num_indexed_dims, device_ptrs=int(device))
    tq, gb, nc = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    table = np.zeros(64, np.float64)
    rc = _L().knnjoin_batch_autotune(*a.call_args(), ctypes.byref(prm),
                                     ctypes.cast(a.pids, ctypes.c_void_p),
                                     ctypes.cast(a.pdists, ctypes.c_void_p), ctypes.byref(tq),
                                     ctypes.byref(gb), _lib.p(table), ctypes.byref(nc))
    _lib.check(rc)
    names = [f"{t}/{g}" for t in (16, 32, 64) for g in (1024, 2048, 4096, 8192)]
    res = {names[i]: float(table[i]) for i in range(nc.value)}
    if cache:
        key = shape_key(a.n, a.d, a.k, num_indexed_dims)
        try:
            with open(cache) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
        db[key] = {"tile_queries": tq.value, "grid_blocks": gb.value, "ms": res}
        with open(cache, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return tq.value, gb.value, res


def shape_key(n, d, k, nd) -> str:
    """Coarse key of a batch shape: set count, total points (log2), dims and k values present."""
    tot = int(np.sum(n))
    return (f"sets{len(n)}_pts2^{int(np.log2(max(tot, 1)))}_d{'-'.join(map(str, sorted(set(map(int, d)))))}"
            f"_k{'-'.join(map(str, sorted(set(map(int, k)))))}_nd{nd}")


def tuned(n, d, k, nd=0, cache: str = TUNE_CACHE):
    """(tile_queries, grid_blocks) stored for this shape key, or (0, 0)."""
    try:
        with open(cache) as f:
            e = json.load(f).get(shape_key(n, d, k, nd))
    except (OSError, ValueError):
        e = None
    return (e["tile_queries"], e["grid_blocks"]) if e else (0, 0)
