"""GPU tests of the 4-D and 6-D workloads (uniform and skewed cubes, k=32, NUMINDEXEDDIM = d
and d-2): plain and learned schedules give the same rows bit for bit, equal to FP64 brute
force on a sample, with the shipped per-dimension radius models when present and with an
untrained model otherwise."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu

CASES = [(4, 4, "uniform"), (4, 2, "skewed"), (6, 6, "uniform"), (6, 4, "skewed")]


def bits(a):
    return a.view(np.uint32)


def _points(d, gen, n, seed):
    return (synthetic.unit_cube if gen == "uniform" else synthetic.skewed_cube)(n, d, seed=seed)


@pytest.mark.parametrize("d,nd,gen", CASES)
def test_plain_and_learned_agree_with_bruteforce(oracle, d, nd, gen):
    loaded = learned_radius.register_default_models()
    if not learned_radius.has_model(loaded, d, nd):
        _lib.set_radius_model(nd, learned_radius.init_weights(d * 10 + nd) * 0.1, 0.3, ndim=d)
    pts = _points(d, gen, 200_000, seed=d * 100 + nd)
    ids_p, d_p, st_p = _lib.knnjoin_ex(pts, 32, num_indexed_dims=nd, beta=0.025)
    ids_l, d_l, st_l = _lib.knnjoin_ex(pts, 32, num_indexed_dims=nd, mode=_lib.KNN_MODE_LEARNED)
    assert np.array_equal(ids_p, ids_l) and np.array_equal(bits(d_p), bits(d_l))
    q = np.random.default_rng(d + nd).choice(len(pts), 300, replace=False)
    perm, _ = oracle.reorder(pts)
    bi, bd = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), 32, queries=q)
    assert np.array_equal(ids_p[q], bi)
    assert np.array_equal(bits(d_p[q]), bits(bd))


def test_model_for_ndim_takes_precedence():
    # a model registered for (NDIM=4, nd=2) is used for 4-D data, the any-NDIM nd=2 model for
    # 3-D data: both give the plain schedule's rows
    pts4 = _points(4, "skewed", 30_000, seed=1)
    pts3 = _points(3, "uniform", 30_000, seed=2)
    if (0, 2) not in learned_radius.register_default_models():
        _lib.set_radius_model(2, learned_radius.init_weights(2) * 0.05, 0.2)
    _lib.set_radius_model(2, learned_radius.init_weights(1) * 0.05, 0.2, ndim=4)
    for pts in (pts4, pts3):
        a = _lib.knnjoin_ex(pts, 16, num_indexed_dims=2, beta=0.025)
        b = _lib.knnjoin_ex(pts, 16, num_indexed_dims=2, mode=_lib.KNN_MODE_LEARNED)
        assert np.array_equal(a[0], b[0]) and np.array_equal(bits(a[1]), bits(b[1]))
