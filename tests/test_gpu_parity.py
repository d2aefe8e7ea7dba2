"""GPU parity tests: the HIP library (through its C ABI) against the CPU oracle and the golden
fixtures.  Bar: bit-exact ids, FP32 distance bits, index arrays and estimator outputs."""
import os

import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


# ---------------------------------------------------------------- stages
@pytest.mark.parametrize("n,d", [(1000, 2), (150000, 3), (20000, 6)])
def test_stage_reorder(oracle, n, d):
    rng = np.random.default_rng(n + d)
    pts = (rng.uniform(0, 1, (n, d)) * rng.uniform(0.5, 2.0, d)).astype(np.float32)
    perm, var = _lib.stage_reorder(pts)
    p2, v2 = oracle.reorder(pts)
    assert list(perm) == list(p2)
    assert np.array_equal(bits(var), bits(v2))


@pytest.mark.parametrize("gen,n,k,beta", [
    ("uniform", 100000, 8, 0.025), ("mixed", 200000, 64, 0.025), ("uniform", 20000, 16, 0.0),
    ("cube4", 60000, 32, 0.025),
])
def test_stage_estimate_eps(oracle, gen, n, k, beta):
    if gen == "uniform":
        pts = synthetic.uniform_box(n)
    elif gen == "mixed":
        pts = synthetic.mixed_box(n)
    else:
        pts = synthetic.unit_cube(n, 4)
    eps, avg, cnt, bucket, hist = _lib.stage_estimate_eps(pts, k, beta)
    rc, info, h2 = oracle.estimate_eps(pts, k, beta)
    assert rc == 0
    assert cnt == info.cnt
    assert avg == info.avg
    assert np.array_equal(hist, h2)
    assert bucket == info.bucket and eps == info.eps


@pytest.mark.parametrize("d,nd,eps", [(2, 2, 0.9), (2, 2, 0.05), (3, 3, 0.07), (4, 2, 0.05), (6, 6, 0.2), (5, 1, 0.01)])
def test_stage_index(oracle, d, nd, eps):
    pts = synthetic.unit_cube(120000, d, seed=d * 100 + nd)
    minArr, nCells, tot, sid, ckeys, cstart = _lib.stage_index(pts, nd, eps)
    mn, mx = oracle.minmax(pts, nd)
    m2, c2, t2 = oracle.grid_dims(mn, mx, eps)
    assert np.array_equal(bits(minArr), bits(m2)) and list(nCells) == list(c2) and tot == t2
    _, sid2, ck2, cs2 = oracle.build_index(pts, nd, eps, m2, c2)
    assert np.array_equal(sid, sid2)
    assert np.array_equal(ckeys, ck2)
    assert np.array_equal(cstart, cs2)


@pytest.mark.parametrize("k,eps", [(8, 2.0), (8, 0.6), (64, 3.0), (130, 4.0)])
def test_stage_join_pass(oracle, k, eps):
    pts = synthetic.uniform_box(30000, seed=k)
    q = np.random.default_rng(k).choice(30000, 5000, replace=False).astype(np.uint32)
    ids, d, failed, nf = _lib.stage_join_pass(pts, 2, k, eps, q)
    i2, d2, f2, nf2 = oracle.join_pass(pts, 2, k, eps, q)
    assert nf == nf2
    assert np.array_equal(failed, f2)
    ok = q[failed[q] == 0]
    assert np.array_equal(ids[ok], i2[ok])
    assert np.array_equal(bits(d[ok]), bits(d2[ok]))


# ---------------------------------------------------------------- whole join
@pytest.mark.parametrize("beta", [0.0, 0.025, 0.2])
def test_join_golden(golden, beta):
    for name, g in golden.items():
        ids, d, st = _lib.knnjoin_ex(g["points"], g["k"], num_indexed_dims=g["nd"], beta=beta)
        assert list(st.perm[: g["points"].shape[1]]) == list(g["perm"]), name
        assert np.array_equal(ids, g["ids"]), name
        assert np.array_equal(bits(d), bits(g["dists"])), name


def _compare_with_oracle(oracle, pts, k, nd=None, beta=0.025):
    ids, d, st = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd or 0, beta=beta)
    nd2 = nd or min(pts.shape[1], 6)
    rc, i2, d2, s2 = oracle.knnjoin(pts, k, nd=nd2, beta=beta)
    assert rc == 0
    assert list(st.perm[: pts.shape[1]]) == list(s2.perm[: pts.shape[1]])
    assert st.eps0 == s2.eps0
    assert st.iterations == s2.iterations and st.reindexes == s2.reindexes
    assert np.array_equal(ids, i2)
    assert np.array_equal(bits(d), bits(d2))
    assert st.checksum == s2.checksum
    return ids, d, st


def test_join_plain_small_workload(oracle):
    # 100,000 points uniform over [0,360)x[-90,90], k=8, BETA=0.025, 2 indexed dims
    pts = synthetic.uniform_box(100000)
    ids, d, st = _compare_with_oracle(oracle, pts, 8)
    sub = np.random.default_rng(1).choice(100000, 300, replace=False)
    bi, bd = oracle.bruteforce(pts, 8, queries=sub)
    assert np.array_equal(ids[sub], bi) and np.array_equal(bits(d[sub]), bits(bd))


@pytest.mark.parametrize("k", [16, 64, 128])
def test_join_mixed_density(oracle, k):
    pts = synthetic.mixed_box(150000, seed=77)
    _compare_with_oracle(oracle, pts, k)


@pytest.mark.parametrize("d,nd,skew", [(3, 3, False), (4, 4, False), (4, 2, True), (6, 6, False), (6, 4, True)])
def test_join_higher_dims(oracle, d, nd, skew):
    pts = (synthetic.skewed_cube if skew else synthetic.unit_cube)(40000, d, seed=d + nd)
    _compare_with_oracle(oracle, pts, 32, nd=nd)


def test_join_double(oracle):
    pts = synthetic.mixed_box(40000, seed=5).astype(np.float64)
    ids, d, st = _lib.knnjoin_ex(pts, 10, beta=0.025)
    rc, i2, d2, s2 = oracle.knnjoin(pts, 10, nd=2, beta=0.025)
    assert rc == 0 and np.array_equal(ids, i2) and np.array_equal(bits(d), bits(d2))


def test_wrapper_flat_interface(oracle):
    from hybrid_k_nearest_neighbor_self_join_amd import knnjoingpu
    pts = synthetic.uniform_box(20000, seed=3)
    ids, d = knnjoingpu.knnjoin(pts.reshape(-1).tolist(), 5, 2, "float", verbose=False)
    rc, i2, d2, _ = oracle.knnjoin(pts, 5, nd=2, beta=0.05)
    assert ids.shape == (20000 * 6,) and ids.dtype == np.int32
    assert np.array_equal(ids, i2.reshape(-1)) and np.array_equal(bits(d), bits(d2.reshape(-1)))
    idsd, dd = knnjoingpu.knnjoin(pts.astype(np.float64).reshape(-1), 5, 2, "double", verbose=True)
    assert dd.dtype == np.float64 and idsd.shape == ids.shape


def test_reference_double_binding_of_the_double_build(oracle):
    # knnjoingpu.py:166-217 with DTYPE="double", unmodified: the library found by name in its
    # directory, KNNJoinPy bound with double ndpointer argtypes, against the DTYPE=double build
    import numpy.ctypeslib as npct
    from ctypes import c_double, c_int, c_uint

    libknnjoin = npct.load_library("libgpuknnselfjoin.so", os.path.dirname(_lib.LIB_DOUBLE_PATH))
    array_1d_double = npct.ndpointer(dtype=c_double, ndim=1, flags="CONTIGUOUS")
    array_1d_int = npct.ndpointer(dtype=c_int, ndim=1, flags="CONTIGUOUS")
    libknnjoin.KNNJoinPy.argtypes = [array_1d_double, c_uint, c_uint, c_uint, array_1d_int, array_1d_double]
    pts = synthetic.mixed_box(30000, seed=6).astype(np.float64)
    KNN, NDIM = 7, 2
    numPoints = len(pts)
    c_DATASET = np.ascontiguousarray(pts.reshape(-1), dtype=c_double)
    ret_outNeighborTable = np.zeros(numPoints * (KNN + 1), dtype=c_int)
    ret_outNeighborTableDistances = np.zeros(numPoints * (KNN + 1), dtype=c_double)
    libknnjoin.KNNJoinPy(c_DATASET, c_uint(numPoints), c_uint(KNN), c_uint(NDIM), ret_outNeighborTable,
                         ret_outNeighborTableDistances)
    rc, i2, d2, _ = oracle.knnjoin(pts, KNN, nd=2, beta=0.05)
    assert rc == 0
    assert np.array_equal(ret_outNeighborTable, i2.reshape(-1))
    assert np.array_equal(bits(ret_outNeighborTableDistances), bits(d2.reshape(-1)))


def test_errors_and_edges(oracle):
    pts = synthetic.uniform_box(10, seed=1)
    with pytest.raises(_lib.KnnError) as e:
        _lib.knnjoin_ex(pts, 10)
    assert e.value.code == _lib.KNN_ERR_TOO_FEW
    # N == K+1 and all-identical points
    ids, d, st = _lib.knnjoin_ex(pts, 9)
    assert np.array_equal(np.sort(ids, 1), np.tile(np.arange(10), (10, 1)))
    ids, d, st = _lib.knnjoin_ex(np.ones((300, 3), np.float32), 4)
    assert np.array_equal(ids, np.tile(np.arange(5, dtype=np.int32), (300, 1))) and np.all(d == 0)
    # N <= 1000: BETA estimate degenerates, bounding-box eps0 is used
    small = synthetic.uniform_box(700, seed=2)
    ids, d, st = _lib.knnjoin_ex(small, 6)
    bi, bd = oracle.bruteforce(small, 6)
    assert np.array_equal(ids, bi) and np.array_equal(bits(d), bits(bd))


@pytest.mark.parametrize("kern", ["collect", "heap"])
def test_join_heavy_duplicates_overflow_kernels(oracle, monkeypatch, kern):
    # 3000 copies of one point inside a uniform background: those queries' staged lists
    # overflow and are finished by the collect kernel (default) or the heap kernel
    monkeypatch.setenv("KNN_OVERFLOW_KERNEL", kern)
    rng = np.random.default_rng(12)
    pts = np.concatenate([np.full((3000, 2), 0.5, np.float32),
                          rng.uniform(0, 1, (20000, 2)).astype(np.float32)])[rng.permutation(23000)]
    ids, d, st = _lib.knnjoin_ex(pts, 16, beta=0.025)
    assert st.overflow_queries > 0
    rc, i2, d2, _ = oracle.knnjoin(pts, 16, nd=2, beta=0.025)
    assert np.array_equal(ids, i2) and np.array_equal(bits(d), bits(d2))


def test_repeated_calls_reuse_buffers(oracle):
    a = synthetic.uniform_box(50000, seed=8)
    b = synthetic.mixed_box(20000, seed=9)
    r1 = _lib.knnjoin_ex(a, 8)
    r2 = _lib.knnjoin_ex(b, 8)
    r3 = _lib.knnjoin_ex(a, 8)
    assert np.array_equal(r1[0], r3[0]) and np.array_equal(bits(r1[1]), bits(r3[1]))
    rc, i2, d2, _ = oracle.knnjoin(b, 8, nd=2, beta=0.05)
    assert np.array_equal(r2[0], i2)


@pytest.mark.parametrize("kernel", ["select", "collect"])
@pytest.mark.parametrize("d,nd,k,gen", [(2, 2, 16, "mixed"), (4, 4, 32, "uniform"), (6, 6, 8, "skewed"),
                                         (6, 4, 16, "uniform"), (3, 3, 200, "uniform")])
def test_both_range_kernels_match_oracle(oracle, monkeypatch, kernel, d, nd, k, gen):
    # the select kernel (two passes, LDS histogram) and the collect kernel (one pass, LDS
    # collector) are chosen by the number of indexed dims; forced here onto the same inputs
    monkeypatch.setenv("KNN_RANGE_KERNEL", kernel)
    if gen == "mixed":
        pts = synthetic.mixed_box(60000, seed=71)
    elif gen == "uniform":
        pts = synthetic.unit_cube(30000, d, seed=72 + d)
    else:
        pts = synthetic.skewed_cube(30000, d, seed=73 + d)
    ids, dist, st = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025)
    rc, i2, d2, _ = oracle.knnjoin(pts, k, nd=nd, beta=0.025)
    assert rc == 0
    assert np.array_equal(ids, i2)
    assert np.array_equal(bits(dist), bits(d2))
