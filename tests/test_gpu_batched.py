"""GPU tests of batched joins: every set's rows equal the single-set join's rows bit for bit
(and FP64 brute force on a sample), across dimensionalities, k values and set sizes, with the
hipGraph captured once and replayed, the eager path, the autotuner and error handling."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, batched, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32)


def _sets(nsets, seed, lo=500, hi=20000):
    return synthetic.batched_sets(nsets, seed=seed, lo=lo, hi=hi)


def _check_against_single(sets, ids, dists):
    for i, (pts, k) in enumerate(sets):
        ri, rd, _ = _lib.knnjoin_ex(pts, k)
        assert np.array_equal(ids[i], ri), f"set {i}: n={len(pts)} d={pts.shape[1]} k={k}"
        assert np.array_equal(bits(dists[i]), bits(rd)), f"set {i}"


def test_batched_equals_single_set_joins():
    sets = _sets(24, 11)
    ids, dists, st = batched.knnjoin_batched([s for s, _ in sets], [k for _, k in sets])
    assert st.groups == len({s.shape[1] for s, _ in sets})
    _check_against_single(sets, ids, dists)


def test_batched_graph_replay_matches_eager_each_time():
    sets = _sets(12, 12)
    pts, ks = [s for s, _ in sets], [k for _, k in sets]
    ref_i, ref_d, st0 = batched.knnjoin_batched(pts, ks, use_graph=False)
    assert st0.graph_replayed == 0
    for rep in range(3):
        ids, dists, st = batched.knnjoin_batched(pts, ks)
        assert st.graph_replayed == (1 if rep else 0)
        for a, b, c, d in zip(ids, ref_i, dists, ref_d):
            assert np.array_equal(a, b) and np.array_equal(bits(c), bits(d))
        # eager work between replays
        _lib.knnjoin_ex(synthetic.uniform_box(5000, seed=rep), 4)
    # new point values, same shapes: the graph is replayed on the new inputs
    rng = np.random.default_rng(5)
    pts2 = [(p + rng.normal(0, 0.01, p.shape)).astype(np.float32) for p in pts]
    ids, dists, st = batched.knnjoin_batched(pts2, ks)
    assert st.graph_replayed == 1
    _check_against_single(list(zip(pts2, ks)), ids, dists)


def test_batched_bruteforce_sample(oracle):
    sets = _sets(6, 13, lo=2000, hi=6000)
    ids, dists, st = batched.knnjoin_batched([s for s, _ in sets], [k for _, k in sets])
    for i, (pts, k) in enumerate(sets):
        q = np.random.default_rng(i).choice(len(pts), 200, replace=False)
        perm, _ = oracle.reorder(pts)
        # distances are summed in the join's (variance) dimension order
        bi, bd = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), k, queries=q)
        assert np.array_equal(ids[i][q], bi)
        assert np.array_equal(bits(dists[i][q]), bits(bd))


@pytest.mark.parametrize("graph_passes", [1, 2, 5])
def test_batched_pass_counts(graph_passes):
    # fewer graph passes leave more to the eager loop; rows never change
    sets = _sets(8, 14)
    pts, ks = [s for s, _ in sets], [k for _, k in sets]
    ids, dists, st = batched.knnjoin_batched(pts, ks, graph_passes=graph_passes)
    assert st.passes_graph == graph_passes
    _check_against_single(sets, ids, dists)


@pytest.mark.parametrize("kern", ["collect", "heap"])
def test_batched_edge_sets(monkeypatch, kern):
    # the staging-overflow queries go to the collect kernel (default) or the heap kernel
    monkeypatch.setenv("KNN_OVERFLOW_KERNEL", kern)
    rng = np.random.default_rng(3)
    sets = [
        (np.ones((40, 3), np.float32), 5),                               # all duplicates
        (rng.uniform(0, 1, (33, 2)).astype(np.float32), 32),             # N == K+1
        (np.zeros((0, 2), np.float32), 8),                               # empty set
        (np.concatenate([np.full((500, 2), 0.5, np.float32),
                         rng.uniform(0, 1, (3000, 2)).astype(np.float32)]), 16),  # heap fallback
        (rng.uniform(0, 1, (4000, 1)).astype(np.float32), 8),            # 1-D
        (rng.uniform(0, 1, (3000, 9)).astype(np.float32), 8),            # 9-D, 6 indexed dims
    ]
    ids, dists, st = batched.knnjoin_batched([s for s, _ in sets], [k for _, k in sets])
    for i, (pts, k) in enumerate(sets):
        if len(pts) == 0:
            assert ids[i].shape == (0, k + 1)
            continue
        ri, rd, _ = _lib.knnjoin_ex(pts, k)
        assert np.array_equal(ids[i], ri), i
        assert np.array_equal(bits(dists[i]), bits(rd)), i


def test_batched_too_few_points_is_an_error():
    sets = [np.zeros((5, 2), np.float32)]
    with pytest.raises(_lib.KnnError) as e:
        batched.knnjoin_batched(sets, [8])
    assert e.value.code == _lib.KNN_ERR_TOO_FEW


def test_autotune_keeps_rows_and_picks_a_candidate(tmp_path):
    sets = _sets(10, 15)
    pts, ks = [s for s, _ in sets], [k for _, k in sets]
    tq, gb, table = batched.autotune(pts, ks, cache=str(tmp_path / "tune.json"))
    assert tq in (16, 32, 64) and gb in (1024, 2048, 4096, 8192) and len(table) == 12
    assert table[f"{tq}/{gb}"] == min(table.values())
    ids, dists, st = batched.knnjoin_batched(pts, ks)
    assert (st.tile_queries, st.grid_blocks) == (tq, gb)
    _check_against_single(sets, ids, dists)
    n = [len(p) for p in pts]
    d = [p.shape[1] for p in pts]
    assert batched.tuned(n, d, ks, cache=str(tmp_path / "tune.json")) == (tq, gb)


def test_batched_256_sets_full_size_against_oracle(oracle):
    # the batched workload at its stated size: 256 sets, sizes log-uniform in [1,000, 200,000],
    # d in {2,3,4,6}, k in {8,32}; captured once, replayed, every set checked against FP64
    # brute force on 300 seeded queries
    sets = synthetic.batched_sets(256)
    pts, ks = [s for s, _ in sets], [k for _, k in sets]
    print(f"batched 256 sets, {sum(len(p) for p in pts)} points", flush=True)
    ids0, dists0, st0 = batched.knnjoin_batched(pts, ks)
    ids, dists, st = batched.knnjoin_batched(pts, ks)
    assert st.graph_replayed == 1 and st.groups == 4
    for i, (p_, k) in enumerate(sets):
        assert np.array_equal(ids[i], ids0[i]) and np.array_equal(bits(dists[i]), bits(dists0[i])), i
        q = np.random.default_rng(1000 + i).choice(len(p_), min(300, len(p_)), replace=False)
        perm, _ = oracle.reorder(p_)
        bi, bd = oracle.bruteforce(np.ascontiguousarray(p_[:, perm]), k, queries=q)
        assert np.array_equal(ids[i][q], bi), f"set {i}: n={len(p_)} d={p_.shape[1]} k={k}"
        assert np.array_equal(bits(dists[i][q]), bits(bd)), f"set {i}"


def test_batched_new_set_shapes_after_graph_replays():
    # a captured graph replayed twice, then sets of other shapes (the plan and its graph are
    # rebuilt), then the first shapes again: every call's rows equal the single-set joins
    a = _sets(16, 21, lo=1000, hi=30000)
    b = _sets(9, 22, lo=2000, hi=40000)
    pa, ka = [s for s, _ in a], [k for _, k in a]
    pb, kb = [s for s, _ in b], [k for _, k in b]
    for rep in range(3):
        ids, dists, st = batched.knnjoin_batched(pa, ka)
        assert st.graph_replayed == (1 if rep else 0)
    _check_against_single(a, ids, dists)
    ids, dists, st = batched.knnjoin_batched(pb, kb)
    assert st.graph_replayed == 0
    _check_against_single(b, ids, dists)
    ids, dists, st = batched.knnjoin_batched(pb, kb)
    assert st.graph_replayed == 1
    _check_against_single(b, ids, dists)
    ids, dists, st = batched.knnjoin_batched(pa, ka)
    assert st.graph_replayed == 0
    _check_against_single(a, ids, dists)
