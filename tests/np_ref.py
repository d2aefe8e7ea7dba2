"""Independent numpy restatements used to pin the C oracle (test code only).

Written separately from oracle/knn_oracle.c so that a slip in one (a dropped term, a wrong
index, a transposed operand) makes the two disagree.  Float32 work uses numpy float32
scalars/arrays, whose +,-,*,/,sqrt are single IEEE operations (numpy does not fuse).
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def reorder_perm(pts: np.ndarray):
    """main.cu:2521-2615 with numpy float32 scalars, sequential sums over rows 0,100,..."""
    n, d = pts.shape
    inv = F32(1.0 / 100.0)
    denom = F32(F32(n) * inv)
    var = []
    for i in range(d):
        col = pts[::100, i].astype(np.float32)
        s = F32(0.0)
        for v in col:
            s = F32(s + v)
        avg = F32(s / denom)
        s = F32(0.0)
        for v in col:
            df = F32(v - avg)
            s = F32(s + F32(df * df))
        var.append(F32(s / denom))
    # stable, descending
    perm = sorted(range(d), key=lambda j: -float(var[j]))
    return np.array(perm, np.uint32), np.array([var[j] for j in perm], np.float32)


def key64(pts_r: np.ndarray, q: int) -> np.ndarray:
    P = pts_r.astype(np.float64)
    dx = P - P[q]
    s = dx[:, 0] * dx[:, 0]
    for j in range(1, P.shape[1]):
        s = s + dx[:, j] * dx[:, j]
    return s


def dist32(pts_r: np.ndarray, q: int, ids: np.ndarray) -> np.ndarray:
    c = pts_r[ids].astype(np.float32)
    df = (c - pts_r[q].astype(np.float32)).astype(np.float32)
    s = (df[:, 0] * df[:, 0]).astype(np.float32)
    for j in range(1, pts_r.shape[1]):
        s = (s + (df[:, j] * df[:, j]).astype(np.float32)).astype(np.float32)
    return np.sqrt(s).astype(np.float32)


def bruteforce(pts_r: np.ndarray, k: int, queries=None):
    """FP64 brute force, rows ordered by (key, id); FP32 fixed-order distances."""
    n = pts_r.shape[0]
    qs = range(n) if queries is None else queries
    ids_all = np.arange(n)
    out_i, out_d = [], []
    for q in qs:
        key = key64(pts_r, q)
        order = np.lexsort((ids_all, key))[: k + 1]
        out_i.append(order.astype(np.int32))
        out_d.append(dist32(pts_r, q, order))
    return np.array(out_i, np.int32), np.array(out_d, np.float32)


def avg_dist(pts_r: np.ndarray):
    """kernel.cu:343-381 restated: sample point g*1000 vs points g%256 + 1000k, summed in the
    fixed order of the GPU and the oracle: segments of L pairs, per-thread sums over the
    segments, sums of 256 consecutive threads, then the total."""
    n, d = pts_r.shape
    tb = max(1, math.ceil((0.001 * n) / 256.0))
    G = tb * 256
    kmax = (n + 999) // 1000
    S = max(1, min((kmax + 63) // 64, max(1, (1 << 26) // G)))
    L = max(1, (kmax + S - 1) // S)
    total = 0.0
    cnt = 0
    for c0 in range(0, G, 256):
        chunk = 0.0
        for g in range(c0, c0 + 256):
            tid = g * 1000
            if tid >= n:
                continue
            idx = np.arange(g % 256, n, 1000)
            df = (pts_r[idx] - pts_r[tid]).astype(np.float32)
            sq = (df * df).astype(np.float32).astype(np.float64)
            run = sq[:, 0].copy()
            for j in range(1, d):
                run = run + sq[:, j]
            dist = np.sqrt(run)
            part = 0.0
            for s0 in range(0, len(dist), L):  # segments of L pairs, each summed in order
                seg = 0.0
                for v in dist[s0:s0 + L]:
                    seg += float(v)
                part += seg
            chunk += part
            cnt += len(idx)
        total += chunk
    return (total / cnt if cnt else 0.0), cnt


def kdist_eps(pts_r: np.ndarray, k: int, beta: float):
    """kernel.cu:450-523 + GPU.cu:3484-3527 restated."""
    n, d = pts_r.shape
    avg, _ = avg_dist(pts_r)
    bw = avg / 10000.0
    if n >= 50000:
        offset = int(1.0 / ((50000.0 / n) * 0.01))
    else:
        offset = int(1.0 / 0.01)
    nsample = n // offset
    hist = np.zeros(10000, np.int64)
    sel = np.concatenate([np.arange(i, min(i + 256, n)) for i in range(0, n, 2048)])
    for b in range(nsample):
        tid = b * offset
        df = (pts_r[sel] - pts_r[tid]).astype(np.float32)
        sq = (df * df).astype(np.float32).astype(np.float64)
        run = sq[:, 0].copy()
        for j in range(1, d):
            run = run + sq[:, j]
        r = np.sqrt(run)
        ok = (r < avg) & (sel != tid)
        bucket = np.floor(r[ok] / bw).astype(np.int64)
        bucket = bucket[bucket < 10000]
        np.add.at(hist, bucket, 8)
    thr = int(k + (k * 100 - k) * beta)
    cum = np.cumsum(hist) / float(nsample)
    hit = np.nonzero(cum > thr)[0]
    i = int(hit[0]) if len(hit) else 9999
    return float(np.float32(bw * (i + 1.0))), i, hist.astype(np.uint32), avg


def cell_coord(x, mn, eps, ncell):
    t = F32(F32(F32(x) - F32(mn)) / F32(eps))
    if not (t > 0):
        return 0
    c = int(t)
    return min(c, ncell - 1)


def grid_dims(mn, mx, eps):
    eps = F32(eps)
    minArr = np.array([F32(F32(v) - eps) for v in mn], np.float32)
    hi = np.array([F32(F32(v) + eps) for v in mx], np.float32)
    nCells = np.array([int(math.ceil(float(F32(F32(h - l) / eps)))) for l, h in zip(minArr, hi)], np.uint32)
    return minArr, nCells, int(np.prod(nCells.astype(np.uint64)))


def build_index(pts_r: np.ndarray, nd: int, eps: float, minArr, nCells):
    n = pts_r.shape[0]
    t = ((pts_r[:, :nd] - minArr[None, :nd]).astype(np.float32) / np.float32(eps)).astype(np.float32)
    t = np.where(t > 0, t, 0)
    c = np.minimum(t.astype(np.int64), nCells[None, :nd].astype(np.int64) - 1)
    mult = np.cumprod(np.concatenate([[1], nCells[:nd - 1].astype(np.int64)]))
    key = (c * mult[None, :]).sum(axis=1).astype(np.uint64)
    order = np.argsort(key, kind="stable").astype(np.uint32)
    sk = key[order]
    first = np.ones(n, bool)
    first[1:] = sk[1:] != sk[:-1]
    starts = np.nonzero(first)[0].astype(np.uint32)
    return key, order, sk[first], np.append(starts, np.uint32(n)).astype(np.uint32)
