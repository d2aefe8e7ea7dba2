"""GPU tests of the hybrid CPU/GPU split (splitWork, main.cu:2733-2933): the queries of the
sparse cells of the first grid are joined by host threads concurrently with the GPU's
failed-query loop over the others; every row equals the oracle's."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def _check(oracle, pts, k, nd=None, **kw):
    d = pts.shape[1]
    nd = nd or min(d, 6)
    ids, dd, st = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025, **kw)
    rc, i2, d2, _ = oracle.knnjoin(pts, k, nd=nd, beta=0.025)
    assert rc == 0
    bad = np.flatnonzero((ids != i2).any(axis=1))
    assert bad.size == 0, f"{bad.size} rows differ, first {bad[0]}"
    assert np.array_equal(bits(dd), bits(d2))
    return st


def test_hybrid_mixed_rows_equal_oracle(oracle):
    pts = synthetic.mixed_box(200_000, seed=41)
    st = _check(oracle, pts, 16, hybrid_threads=8)
    assert st.hybrid_queries > 0 and st.failed_final == 0
    st2 = _check(oracle, pts, 16, hybrid_threads=4, hybrid_max_queries=1000)
    assert 0 < st2.hybrid_queries <= 1000


def test_hybrid_far_outliers_and_projection(oracle):
    rng = np.random.default_rng(0x5EED0300)
    pts = np.concatenate([rng.normal(0.0, 1.0, (60_000, 2)), rng.uniform(30.0, 40.0, (20, 2))]).astype(np.float32)
    st = _check(oracle, pts, 16, hybrid_threads=8, brute_fallback=0)
    assert st.hybrid_queries >= 20
    sk = synthetic.skewed_cube(60_000, 4, seed=42)
    st = _check(oracle, sk, 32, nd=2, hybrid_threads=8)
    assert st.hybrid_queries > 0


def test_hybrid_double_and_edges(oracle):
    pts = synthetic.mixed_box(30_000, seed=43).astype(np.float64)
    _check(oracle, pts, 10, hybrid_threads=3)
    ones = np.ones((300, 3), np.float32)  # one cell: nothing sparse
    ids, d, st = _lib.knnjoin_ex(ones, 4, hybrid_threads=4)
    assert np.array_equal(ids, np.tile(np.arange(5, dtype=np.int32), (300, 1)))


def test_hybrid_needs_the_beta_schedule():
    pts = synthetic.mixed_box(5000, seed=44)
    with pytest.raises(_lib.KnnError) as e:
        _lib.knnjoin_ex(pts, 8, mode=_lib.KNN_MODE_LEARNED, hybrid_threads=4)
    assert e.value.code == _lib.KNN_ERR_ARG
