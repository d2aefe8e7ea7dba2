"""Benchmark: exact KNN self-join throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--k 64] [--n 10000000] [--mode auto]

Workload (default): 10,000,000 2-D float32 points, 70% uniform over [0,360)x[-90,90] and 30%
from 64 Gaussian clusters (sigma 0.5), k=64, BETA=0.025, 2 indexed dims.  One step = the
whole device pipeline of one self-join (dimension reorder, radius schedule, grid index,
every failed-query pass, checksum) from HBM-resident points to the HBM-resident
(K+1)-table.  Both radius schedules are timed when a trained 2-D radius model ships with the
package: the plain BETA schedule and the learned per-query radius; value = queries/s
(N / time per step) of the faster one, with the learned/plain speedup beside it.

Also reported: the range kernel's achieved algorithmic HBM bandwidth against the 8 TB/s peak
(HIP events on the kernel's stream, inside the library), its measured HBM traffic per launch
when a rocprofv3 counter measurement of this workload is committed, and a CPU baseline: the
oracle's OpenMP restatement of the same grid join on a bounded sample of the same point set.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(pts: np.ndarray, k: int, beta: float, mode: int, steps: int, warmup: int,
            nd: int = 0, keep_rows: bool = False):
    import torch

    from hybrid_k_nearest_neighbor_self_join_amd import _lib

    L = _lib.lib()
    n, d = pts.shape
    dev = torch.device("cuda:0")
    dpts = torch.from_numpy(pts).to(dev)
    ids = torch.empty((n, k + 1), dtype=torch.int32, device=dev)
    dists = torch.empty((n, k + 1), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev)
    prm = _lib.default_params(beta=beta, mode=mode, device_ptrs=1, num_indexed_dims=nd,
                              stream=ctypes.c_void_p(stream.cuda_stream).value)
    torch.cuda.synchronize()

    def step():
        st = _lib.KnnStats()
        rc = L.knnjoin_ex(ctypes.c_void_p(dpts.data_ptr()), _lib.KNN_FLOAT32, n, d, k,
                          ctypes.byref(prm), ctypes.c_void_p(ids.data_ptr()),
                          ctypes.c_void_p(dists.data_ptr()), ctypes.byref(st))
        _lib.check(rc)
        return st

    for i in range(warmup):
        st = step()
        log(f"[bench] warmup {i}: {st.ms_total:.2f} ms, iterations {st.iterations}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(steps):
        stats.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    last = stats[-1]
    out = {
        "ms_per_step": wall * 1e3 / steps,
        "ms_device_median": float(np.median([s.ms_total for s in stats])),
        "kernel_ms": sum(s.ms_range_kernel for s in stats),
        "kernel_launches": sum(s.kernel_launches for s in stats),
        "algo_bytes": sum(s.algo_bytes for s in stats),
        "stats": last.as_dict(),
        "ids_row0": ids[0].cpu().numpy().tolist()[:4],
    }
    if keep_rows:
        import hashlib
        h = hashlib.sha1(ids.cpu().numpy().tobytes())
        h.update(dists.cpu().numpy().tobytes())
        out["rows_digest"] = h.hexdigest()
        ids_host = ids.cpu().numpy()
        out["ids_sample"] = lambda q: ids_host[q]
    return out


def cpu_baseline(n: int, k: int, beta: float, queries: int):
    """The oracle's OpenMP grid join on the host cores, in a child process started with
    OMP_PROC_BIND=close and OMP_NUM_THREADS = the cores used (oracle/cpu_baseline.py: a
    seeded subset of `queries` queries against all n points of the same set, 1 warm-up +
    median of 3, projected to the full join)."""
    import subprocess

    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_NUM_THREADS=str(cores))
    cmd = [sys.executable, "-m", "oracle.cpu_baseline", "--n", str(n), "--k", str(k),
           "--beta", str(beta), "--queries", str(queries), "--threads", str(cores)]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        log(f"[bench] cpu baseline failed: {r.stderr[-2000:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def load_traffic(workload_key: str):
    """HBM bytes per launch of the range kernel measured with rocprofv3 PMC counters
    (FETCH_SIZE / WRITE_SIZE, scaled by the same counters' reading on a copy of known size),
    committed beside the profile summaries; None when no measurement of this workload exists."""
    path = os.path.join(ROOT, "profiles", "k_select_traffic.json")
    try:
        with open(path) as f:
            e = json.load(f).get(workload_key)
    except (OSError, ValueError):
        return None
    return e


def load_counters(workload_key: str):
    """The range kernel's issue counters for this workload from a committed rocprofv3 --pmc
    measurement (SQ_INSTS_VALU, SQ_WAVES, SQ_BUSY_CYCLES, ...): VALU wave instructions per
    launch, the fraction of the VALU issue peak they use at the traced launch time, and VALU
    lane instructions per (query, candidate) pair; None when no measurement exists."""
    path = os.path.join(ROOT, "profiles", "k_select_counters.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload_key)
    except (OSError, ValueError):
        return None


def suite_highdim(args):
    """5,000,000 points in 4-D and 6-D (uniform and skewed cubes), k=32, NUMINDEXEDDIM = d and
    d-2: plain vs learned schedule, one JSON line per case."""
    import torch  # noqa: F401  (before the library: one HIP runtime per process)

    from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic
    from oracle import oracle

    models = learned_radius.register_default_models()
    n = args.n if args.n != 10_000_000 else 5_000_000
    cases = [(4, 4, "uniform"), (4, 2, "uniform"), (4, 4, "skewed"), (4, 2, "skewed"),
             (6, 6, "uniform"), (6, 4, "uniform"), (6, 6, "skewed"), (6, 4, "skewed")]
    if args.case:
        d0, nd0, gen0 = args.case.split(",")
        cases = [(int(d0), int(nd0), gen0)]
    for d, nd, gen in cases:
        pts = (synthetic.unit_cube if gen == "uniform" else synthetic.skewed_cube)(n, d, seed=0x5EED0100 + d * 10 + nd)
        out = {"suite": "highdim", "n": n, "d": d, "nd": nd, "data": gen, "k": 32}
        res = {}
        for mode in (["plain", "learned"] if learned_radius.has_model(models, d, nd) else ["plain"]):
            m = _lib.KNN_MODE_LEARNED if mode == "learned" else _lib.KNN_MODE_BETA
            g = run_gpu(pts, 32, 0.025, m, args.steps, args.warmup, nd=nd, keep_rows=True)
            res[mode] = g
            kms = g["stats"]["ms_range_kernel"]
            out[mode] = {"queries_per_s": n / (g["ms_per_step"] / 1e3), "ms_per_step": g["ms_per_step"],
                         "iterations": g["stats"]["iterations"], "candidates": g["stats"]["candidates"],
                         "range_kernel_ms": kms,
                         "candidates_per_s": g["stats"]["candidates"] / (kms / 1e3) if kms > 0 else None,
                         "checksum": g["stats"]["checksum"]}
        if "learned" in res:
            out["speedup_learned_vs_plain"] = res["plain"]["ms_per_step"] / res["learned"]["ms_per_step"]
            out["identical_rows"] = bool(res["plain"]["rows_digest"] == res["learned"]["rows_digest"])
        q = np.random.default_rng(d * nd).choice(n, 500, replace=False)
        perm, _ = oracle.reorder(pts)
        bi, _ = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), 32, queries=q)
        out["bruteforce_sample_match"] = bool(np.array_equal(res["plain"]["ids_sample"](q), bi))
        print(json.dumps(out), flush=True)


def suite_batched(args):
    """256 sets, sizes log-uniform in [1,000, 200,000], d in {2,3,4,6}, k in {8,32}: batched
    join (hipGraph replay, autotuned) vs the same sets joined one by one."""
    import torch  # noqa: F401

    from hybrid_k_nearest_neighbor_self_join_amd import _lib, batched, synthetic

    sets = synthetic.batched_sets(256)
    pts, ks = [s for s, _ in sets], [k for _, k in sets]
    N = sum(len(p) for p in pts)
    batched.knnjoin_batched(pts, ks)  # build + capture
    tq, gb, table = batched.autotune(pts, ks, cache=None)
    times = []
    for _ in range(max(args.steps, 3)):
        ids, dists, st = batched.knnjoin_batched(pts, ks)
        times.append(st.ms_device)
    if args.batched_only:
        print(json.dumps({"suite": "batched", "sets": len(sets), "points": N,
                          "batched": {"ms_device_median": float(np.median(times)),
                                      "queries_per_s": N / (np.median(times) / 1e3),
                                      "tile_queries": tq, "grid_blocks": gb,
                                      "graph_replayed": st.graph_replayed}}), flush=True)
        return
    lat = []
    mism = 0
    for i, (p_, k) in enumerate(sets):
        ri, rd, s1 = _lib.knnjoin_ex(p_, k, beta=0.025)
        lat.append(s1.ms_total)
        mism += int(not (np.array_equal(ri, ids[i]) and np.array_equal(rd.view(np.uint32), dists[i].view(np.uint32))))
    lat = np.array(lat)
    print(json.dumps({
        "suite": "batched", "sets": len(sets), "points": N,
        "batched": {"ms_device_median": float(np.median(times)), "queries_per_s": N / (np.median(times) / 1e3),
                    "tile_queries": tq, "grid_blocks": gb, "graph_replayed": st.graph_replayed,
                    "autotune_ms": table},
        "one_by_one": {"ms_total": float(lat.sum()), "queries_per_s": N / (lat.sum() / 1e3),
                       "per_set_ms_p50": float(np.percentile(lat, 50)),
                       "per_set_ms_p99": float(np.percentile(lat, 99))},
        "speedup_batched_vs_one_by_one": float(lat.sum() / np.median(times)),
        "rows_differing_from_single_set_joins": mism,
    }), flush=True)


def suite_brute(args):
    """Brute-force path: whole joins by KNN_MODE_BRUTE against the grid schedule (2-D uniform,
    k 8 and 64), and a dense cluster with 20 far outliers, where the plain loop needs hundreds
    of index builds, with the brute-force fallback off and on."""
    import torch  # noqa: F401

    from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

    learned_radius.register_default_models()

    def best(pts, k, **kw):
        _lib.knnjoin_ex(pts, k, **kw)
        runs = [_lib.knnjoin_ex(pts, k, **kw) for _ in range(max(args.steps, 3))]
        return min(runs, key=lambda r: r[2].ms_total)

    for n in (20_000, 100_000, 300_000):
        pts = synthetic.uniform_box(n, seed=0x5EED0200 + n)
        for k in (8, 64):
            ib, db, sb = best(pts, k, mode=_lib.KNN_MODE_BRUTE)
            ip, dp, sp = best(pts, k, beta=0.025, brute_fallback=0)
            print(json.dumps({
                "suite": "brute", "case": "whole_join", "n": n, "d": 2, "k": k,
                "brute_ms": sb.ms_total, "brute_kernel_ms": sb.ms_join,
                "pairs_per_s": n * n / (sb.ms_join / 1e3), "plain_ms": sp.ms_total,
                "identical_rows": bool(np.array_equal(ib, ip) and np.array_equal(db.view(np.uint32), dp.view(np.uint32))),
            }), flush=True)
    rng = np.random.default_rng(0x5EED0300)
    pts = np.concatenate([rng.normal(0.0, 1.0, (60_000, 2)), rng.uniform(30.0, 40.0, (20, 2))]).astype(np.float32)
    res = {}
    for tag, kw in [("plain_no_fallback", dict(beta=0.025, brute_fallback=0)),
                    ("plain_fallback", dict(beta=0.025, brute_fallback=1)),
                    ("learned_no_fallback", dict(mode=_lib.KNN_MODE_LEARNED, brute_fallback=0)),
                    ("learned_fallback", dict(mode=_lib.KNN_MODE_LEARNED, brute_fallback=1))]:
        if "learned" in tag and not learned_radius.has_model(learned_radius.register_default_models(), 2, 2):
            continue
        i, d, st = best(pts, 16, **kw)
        res[tag] = (i, d, st)
    base = res["plain_no_fallback"]
    print(json.dumps({
        "suite": "brute", "case": "cluster_60000_plus_20_far_outliers", "k": 16,
        **{t: {"ms": r[2].ms_total, "iterations": r[2].iterations, "reindexes": r[2].reindexes,
               "brute_queries": r[2].brute_queries} for t, r in res.items()},
        "identical_rows": all(np.array_equal(r[0], base[0]) and np.array_equal(r[1].view(np.uint32), base[1].view(np.uint32))
                              for r in res.values()),
    }), flush=True)


def suite_hybrid(args):
    """Hybrid CPU/GPU split (splitWork, main.cu:2733-2933) against the GPU-only BETA schedule:
    the far-outlier cluster, 5M skewed 4-D points indexed on 2 dims, and the 10M headline set
    (host share capped), each with the host thread count of the box.  One JSON line per case
    with both times, the host share and whether the rows are identical."""
    import torch  # noqa: F401

    from hybrid_k_nearest_neighbor_self_join_amd import _lib, synthetic

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    rng = np.random.default_rng(0x5EED0300)
    cases = [
        ("cluster_60000_plus_20_far_outliers", np.concatenate(
            [rng.normal(0.0, 1.0, (60_000, 2)), rng.uniform(30.0, 40.0, (20, 2))]).astype(np.float32), 16, 2, 0),
        ("skewed_4d_nd2_5M", synthetic.skewed_cube(5_000_000, 4, seed=0x5EED0100 + 42), 32, 2, 0),
        ("mixed_2d_10M_k64_host_share_1e5", synthetic.mixed_box(10_000_000), 64, 2, 100_000),
        ("mixed_2d_10M_k64_host_share_1e6", synthetic.mixed_box(10_000_000), 64, 2, 1_000_000),
    ]
    for tag, pts, k, nd, cap in cases:
        def best(**kw):
            _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025, brute_fallback=0, **kw)
            runs = [_lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025, brute_fallback=0, **kw)
                    for _ in range(max(args.steps, 2))]
            return min(runs, key=lambda r: r[2].ms_total)
        ig, dg, sg = best()
        ih, dh, sh = best(hybrid_threads=threads, hybrid_max_queries=cap)
        print(json.dumps({
            "suite": "hybrid", "case": tag, "n": len(pts), "d": pts.shape[1], "nd": nd, "k": k,
            "host_threads": threads, "host_share_cap": cap,
            "gpu_only_ms": sg.ms_total, "gpu_only_iterations": sg.iterations, "gpu_only_reindexes": sg.reindexes,
            "hybrid_ms": sh.ms_total, "hybrid_host_queries": sh.hybrid_queries,
            "hybrid_host_ms": sh.ms_hybrid_host, "hybrid_gpu_iterations": sh.iterations,
            "hybrid_vs_gpu_only": sg.ms_total / sh.ms_total,
            "identical_rows": bool(np.array_equal(ig, ih) and np.array_equal(dg.view(np.uint32), dh.view(np.uint32))),
        }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--beta", type=float, default=0.025)
    ap.add_argument("--mode", choices=["auto", "plain", "learned"], default="auto",
                    help="auto: time the plain BETA schedule and, when a trained radius model "
                         "ships with the package, the learned schedule; value = the faster")
    ap.add_argument("--cpu-queries", type=int, default=100_000,
                    help="queries of the CPU baseline's subset join against the full set")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--suite", choices=["highdim", "batched", "brute", "hybrid"], default=None,
                    help="instead of the headline line: the 4-D/6-D cases, the 256-set batch, or "
                         "the brute-force path")
    ap.add_argument("--case", default=None,
                    help="--suite highdim: one case only, as d,nd,uniform|skewed (for profiling)")
    ap.add_argument("--batched-only", action="store_true",
                    help="--suite batched: time the batched join only (no one-by-one comparison), "
                         "for profiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 1 GiB (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": a.numel() * 4,
                          "copies": 5}), flush=True)
        return

    if args.suite == "highdim":
        return suite_highdim(args)
    if args.suite == "batched":
        return suite_batched(args)
    if args.suite == "brute":
        return suite_brute(args)
    if args.suite == "hybrid":
        return suite_hybrid(args)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    # torch ships its own HIP runtime under the same soname as ROCm's: it must be loaded before
    # the library, or the process ends up with two runtimes
    import torch  # noqa: F401

    from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

    t0 = time.perf_counter()
    pts = synthetic.mixed_box(args.n)
    log(f"[bench] generated {args.n} points in {time.perf_counter() - t0:.1f} s")
    models = learned_radius.register_default_models()
    if args.mode == "plain":
        modes = ["plain"]
    elif args.mode == "learned":
        if not learned_radius.has_model(models, 2, 2):
            print("no trained 2-D radius model in the package", file=sys.stderr)
            sys.exit(2)
        modes = ["learned"]
    else:
        modes = ["plain"] + (["learned"] if learned_radius.has_model(models, 2, 2) else [])
    res = {}
    for m in modes:
        mode = _lib.KNN_MODE_LEARNED if m == "learned" else _lib.KNN_MODE_BETA
        g = run_gpu(pts, args.k, args.beta, mode, args.steps, args.warmup)
        g["qps"] = args.n / (g["ms_per_step"] / 1e3)
        res[m] = g
        log(f"[bench] {m}: {g['ms_per_step']:.2f} ms/step, {g['qps'] / 1e6:.2f} M queries/s")
    best = max(res, key=lambda m: res[m]["qps"])
    g = res[best]
    st = g["stats"]
    kern_s = g["kernel_ms"] / 1e3
    achieved = (g["algo_bytes"] / kern_s / 1e9) if kern_s > 0 else 0.0
    launches = max(g["kernel_launches"], 1)
    wkey = f"mixed{args.n}_k{args.k}_{best}"
    tr = load_traffic(wkey)
    line = {
        "metric": "queries/s, 10M 2-D float32 k=64 exact; learned-radius vs plain BETA speedup",
        "value": g["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded mixed-density point set, generated on the host, resident in HBM)",
        "config": {
            "workload": (f"exact KNN self-join, {args.n} 2-D float32 points (70% uniform box + 30% "
                         f"in 64 Gaussian clusters), k={args.k}, BETA={args.beta}, "
                         f"{best} radius schedule"),
            "n_points": args.n, "dim": 2, "k": args.k, "beta": args.beta, "mode": best,
            "num_indexed_dims": 2,
        },
        "modes": {m: {"queries_per_s": r["qps"], "ms_per_step": r["ms_per_step"],
                      "ms_reference_total": r["stats"]["ms_reference_total"],
                      "iterations": r["stats"]["iterations"], "reindexes": r["stats"]["reindexes"],
                      "host_syncs": r["stats"]["host_syncs"],
                      "candidates": r["stats"]["candidates"],
                      "range_kernel_ms": r["stats"]["ms_range_kernel"],
                      "checksum": r["stats"]["checksum"]} for m, r in res.items()},
        "speedup_learned_vs_plain": (res["learned"]["qps"] / res["plain"]["qps"]
                                     if "learned" in res and "plain" in res else None),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": tr["bytes_per_launch"] if tr else None,
            "kernel": "k_select",
            "algo_bytes_per_launch": g["algo_bytes"] / launches,
            "avg_launch_ms": g["kernel_ms"] / launches,
            "launches_per_step": g["kernel_launches"] / max(args.steps, 1),
            "traffic_source": tr["source"] if tr else None,
            # what binds instead of HBM: VALU issue (peak 256 CU x 4 SIMD x 2.4 GHz / 4 cycles
            # per wave64 instruction = 614 G wave instructions/s) and its latency
            "valu": load_counters(wkey),
        },
        "pipeline": {
            "ms_device_median": g["ms_device_median"],
            # the reference's totalTime (main.cu:440-445): the pipeline less the first index build
            "ms_reference_total": st["ms_reference_total"], "ms_index_first": st["ms_index_first"],
            "host_syncs": st["host_syncs"],
            "iterations": st["iterations"], "reindexes": st["reindexes"],
            "eps0": st["eps0"], "eps_final": st["eps_final"],
            "ms_reorder": st["ms_reorder"], "ms_eps": st["ms_eps"], "ms_index": st["ms_index"],
            "ms_join": st["ms_join"], "ms_range_kernel": st["ms_range_kernel"],
            "candidates": st["candidates"], "checksum": st["checksum"],
        },
    }
    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_queries} queries against {args.n} points ...")
        line["cpu_baseline"] = cpu_baseline(args.n, args.k, args.beta, args.cpu_queries)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
