"""CPU baseline of the headline workload: the oracle's OpenMP grid join (TEST INFRASTRUCTURE).

    python -m oracle.cpu_baseline --n 10000000 --k 64 --queries 100000 --threads 16

Run by bench.py as a child process, so that OMP_PROC_BIND / OMP_NUM_THREADS are in place
before the OpenMP runtime starts.  It joins a seeded random subset of `--queries` queries
against the full point set with the plain method (dimension reorder, BETA epsilon, grid,
failed-query loop: the same per-query work as the GPU's full join, every index built over
all N points), one warm-up and then `--repeats` timed runs.  A full join runs the same
set-up and index builds once and the passes over N queries instead of the subset's, so its
time is projected as  t_prep + t_index + t_passes * N / queries  from the median run, and
the reported value is N / that time.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--beta", type=float, default=0.025)
    ap.add_argument("--queries", type=int, default=100_000)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=0x5EED0002)
    args = ap.parse_args()

    from hybrid_k_nearest_neighbor_self_join_amd import synthetic
    from oracle import oracle

    threads = args.threads or int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    pts = synthetic.mixed_box(args.n, seed=args.seed)
    rng = np.random.default_rng(0x5EED00C0)
    q = np.sort(rng.choice(args.n, min(args.queries, args.n), replace=False)).astype(np.uint32)

    def run():
        t0 = time.perf_counter()
        rc, _, _, st = oracle.knnjoin(pts, args.k, nd=2, beta=args.beta, nthreads=threads, subset=q)
        wall = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit(f"oracle join failed: rc={rc}")
        return wall, st.ms_prep / 1e3, st.ms_index / 1e3, st.ms_passes / 1e3, st.iterations, st.reindexes

    run()  # warm-up
    runs = sorted((run() for _ in range(args.repeats)), key=lambda r: r[0])
    wall, prep, index, passes, iters, reidx = runs[len(runs) // 2]
    full = prep + index + passes * (args.n / len(q))
    print(json.dumps({
        "value": args.n / full,
        "unit": "queries/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle plain-method grid join (OpenMP, {threads} threads, "
                   f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', 'unset')}) of {len(q)} seeded "
                   f"random queries against all {args.n} points of the headline set, k={args.k}, "
                   f"BETA={args.beta}; median of {args.repeats} after 1 warm-up: {wall:.2f} s "
                   f"(set-up {prep:.2f} s, {reidx + 1} index builds {index:.2f} s, {iters} passes "
                   f"{passes:.2f} s); value = N / (set-up + index builds + passes x N/queries) = "
                   f"{full:.1f} s projected for the full join"),
        "subset_queries_per_s": len(q) / wall,
        "projected_full_join_s": full,
    }), flush=True)


if __name__ == "__main__":
    main()
