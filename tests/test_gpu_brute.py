"""Brute-force path (the reference's no-selectivity fallback, distanceTableNDBruteForce
GPU.cu:1903-2930 / kernelNDBruteForce kernel.cu:902-975, chosen by checkIndexSelectivity
main.cu:2687-2716): KNN_MODE_BRUTE for whole joins, and the fallback inside the plain and
learned loops.  Bar: the same rows as the oracle's join (ids exact, distance bits exact)."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def assert_rows(oracle, pts, k, ids, d, nd=None):
    rc, i2, d2, _ = oracle.knnjoin(pts, k, nd=nd, beta=0.05)
    assert rc == 0
    assert np.array_equal(ids, i2)
    assert np.array_equal(bits(d), bits(d2))


# sizes: one chunk per query tile (n >= 32768 queries), several chunks (partial tables + merge)
@pytest.mark.parametrize("n,d,k,dtype", [
    (40000, 2, 8, np.float32),     # one chunk, written directly
    (20000, 2, 64, np.float32),    # two chunks per tile, merge kernel
    (5000, 3, 16, np.float32),     # five chunks
    (3000, 6, 32, np.float64),
    (1500, 16, 5, np.float32),
    (2000, 1, 7, np.float64),
])
def test_brute_mode_matches_oracle(oracle, n, d, k, dtype):
    pts = synthetic.unit_cube(n, d, seed=n + d).astype(dtype)
    ids, dist, st = _lib.knnjoin_ex(pts, k, mode=_lib.KNN_MODE_BRUTE)
    assert st.brute_queries == n and st.iterations == 1
    assert_rows(oracle, pts, k, ids, dist, nd=min(d, 6))


def test_brute_large_k(oracle):
    # K+1 = 301: one query per wave (LDS), several chunks
    pts = synthetic.mixed_box(6000, seed=3)
    ids, dist, st = _lib.knnjoin_ex(pts, 300, mode=_lib.KNN_MODE_BRUTE)
    assert_rows(oracle, pts, 300, ids, dist)


def test_brute_duplicates_and_exact_size(oracle):
    # every point identical: rows ranked by id (self is not necessarily first)
    same = np.ones((300, 3), np.float32)
    ids, dist, st = _lib.knnjoin_ex(same, 4, mode=_lib.KNN_MODE_BRUTE)
    assert np.array_equal(ids, np.tile(np.arange(5, dtype=np.int32), (300, 1)))
    assert not dist.any()
    # N = K+1: every row holds every point
    small = synthetic.uniform_box(17, seed=9)
    ids, dist, st = _lib.knnjoin_ex(small, 16, mode=_lib.KNN_MODE_BRUTE)
    assert_rows(oracle, small, 16, ids, dist)
    assert all(sorted(r) == list(range(17)) for r in ids)
    # lattice with many equal distances
    g = np.stack(np.meshgrid(np.arange(40), np.arange(40)), -1).reshape(-1, 2).astype(np.float32)
    ids, dist, st = _lib.knnjoin_ex(g, 12, mode=_lib.KNN_MODE_BRUTE)
    assert_rows(oracle, g, 12, ids, dist)


def test_brute_too_large_k_is_an_error():
    pts = synthetic.uniform_box(3000, seed=1)
    with pytest.raises(RuntimeError, match="KNN_ERR_ARG"):
        _lib.knnjoin_ex(pts, 2048, mode=_lib.KNN_MODE_BRUTE)


def far_outliers(n, seed):
    # a dense cluster plus a few points far away: the plain loop grows eps by eps0/2 per step,
    # so the outliers need a hundred or more index builds; the fallback serves them by brute
    # force
    rng = np.random.default_rng(seed)
    c = rng.normal(0.0, 1.0, (n, 2))
    far = rng.uniform(30.0, 40.0, (20, 2))
    return np.concatenate([c, far]).astype(np.float32)


def test_plain_fallback_same_rows(oracle):
    pts = far_outliers(60000, 4)
    a = _lib.knnjoin_ex(pts, 16, beta=0.025, brute_fallback=0)
    b = _lib.knnjoin_ex(pts, 16, beta=0.025, brute_fallback=1)
    assert a[2].brute_queries == 0 and b[2].brute_queries > 0
    assert b[2].iterations < a[2].iterations
    assert np.array_equal(a[0], b[0]) and np.array_equal(bits(a[1]), bits(b[1]))
    assert_rows(oracle, pts, 16, b[0], b[1])


def test_whole_join_small_n_fallback(oracle):
    # N*N <= 2^28: with the fallback on, the whole join is one brute-force pass
    pts = synthetic.mixed_box(12000, seed=8)
    ids, dist, st = _lib.knnjoin_ex(pts, 10, beta=0.025, brute_fallback=1)
    assert st.brute_queries == 12000
    assert_rows(oracle, pts, 10, ids, dist)


def test_learned_fallback_default(oracle):
    if not learned_radius.has_model(learned_radius.register_default_models(), 2, 2):
        pytest.skip("no 2-D radius model")
    pts = far_outliers(80000, 6)
    a = _lib.knnjoin_ex(pts, 32, mode=_lib.KNN_MODE_LEARNED)
    b = _lib.knnjoin_ex(pts, 32, mode=_lib.KNN_MODE_LEARNED, brute_fallback=0)
    assert a[2].brute_queries > 0 and b[2].brute_queries == 0
    assert np.array_equal(a[0], b[0]) and np.array_equal(bits(a[1]), bits(b[1]))
    assert_rows(oracle, pts, 32, a[0], a[1])
