"""Outlier scores (printOutlierScores, main.cu:952-1067): the oracle against a numpy
restatement and a hand-computed case (CPU), and the GPU entry point against the oracle on a
join's table and on tables with many ties (GPU)."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, synthetic


def np_outlier(ids, dists):
    n, k1 = ids.shape
    s = np.zeros(n, np.float32)
    for e in range(k1):  # FP32, entry order
        s = (s + dists[:, e]).astype(np.float32)
    deg = np.bincount(ids.reshape(-1), minlength=n).astype(np.uint32)
    order = np.lexsort((np.arange(n), -s.astype(np.float64)))
    dr = np.empty(n, np.uint32)
    dr[order] = np.arange(n)
    order = np.lexsort((np.arange(n), deg))
    gr = np.empty(n, np.uint32)
    gr[order] = np.arange(n)
    return s, dr, deg, gr


def test_oracle_hand_computed(oracle):
    # 4 points, k1 = 2: rows (self, nearest)
    ids = np.array([[0, 1], [1, 0], [2, 1], [3, 2]], np.int32)
    d = np.array([[0, 1.0], [0, 1.0], [0, 2.0], [0, 5.0]], np.float32)
    s, dr, deg, gr = oracle.outlier_scores(ids, d)
    assert list(s) == [1.0, 1.0, 2.0, 5.0]
    assert list(dr) == [2, 3, 1, 0]          # 5 > 2 > 1 = 1 (ties by id)
    assert list(deg) == [2, 3, 2, 1]
    assert list(gr) == [1, 3, 2, 0]          # 1 < 2 = 2 < 3


def test_oracle_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    for n, k1 in [(1000, 9), (5000, 33), (300, 2)]:
        ids = rng.integers(0, n, (n, k1)).astype(np.int32)
        ids[:, 0] = np.arange(n)
        d = np.round(rng.uniform(0, 4, (n, k1)), 1).astype(np.float32)  # many ties
        d[:, 0] = 0
        got = oracle.outlier_scores(ids, d)
        want = np_outlier(ids, d)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


@pytest.mark.gpu
def test_gpu_matches_oracle_on_join_table(oracle):
    pts = synthetic.mixed_box(100000, seed=41)
    ids, d, _ = _lib.knnjoin_ex(pts, 8, beta=0.025)
    got = _lib.outlier_scores(ids, d)
    want = oracle.outlier_scores(ids, d)
    assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
    for g, w in zip(got[1:], want[1:]):
        assert np.array_equal(g, w)


@pytest.mark.gpu
def test_gpu_ties_ranked_by_id(oracle):
    rng = np.random.default_rng(5)
    n, k1 = 20000, 17
    ids = rng.integers(0, n, (n, k1)).astype(np.int32)
    d = np.round(rng.uniform(0, 2, (n, k1)), 1).astype(np.float32)
    got = _lib.outlier_scores(ids, d)
    want = oracle.outlier_scores(ids, d)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
