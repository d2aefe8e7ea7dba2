"""GPU tests of the slab order: with fewer indexed than data dims, the points of each cell are
sorted by the quantised first non-indexed coordinate and the select kernel cuts every cell of a
block to the query's (or the tile's) slab range by binary searches.  Only the candidate streams
change, never the rows: the join equals the oracle's, and equals the join with the slab order
turned off (KNN_SLAB=0), bit for bit, for every number of indexed dims the select kernel builds
(1 to 4), 6 indexed dims on the collect kernel, lattice data full of ties, and a non-indexed
dim of zero range."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32)


def _points(d, gen, n, seed):
    if gen == "uniform":
        return synthetic.unit_cube(n, d, seed=seed)
    if gen == "skewed":
        return synthetic.skewed_cube(n, d, seed=seed)
    if gen == "lattice":
        # coordinates on a 1/16 lattice: many equal slab codes and many equal points
        return (np.round(synthetic.unit_cube(n, d, seed=seed) * 16.0) / 16.0).astype(np.float32)
    if gen == "flat":
        # the first non-indexed dim after the variance reorder is constant: no slab order
        p = synthetic.unit_cube(n, d, seed=seed)
        p[:, -1] = 0.25
        return p
    raise ValueError(gen)


# (8, 6) and (7, 6) run on the collect kernel (6 indexed dims, the reference's default above
# 6 data dims), the others on the select kernel
ORACLE_CASES = [(3, 1, "uniform", 16), (4, 2, "skewed", 16), (5, 3, "uniform", 8), (6, 4, "uniform", 16),
                (4, 2, "lattice", 16), (4, 3, "flat", 8), (8, 6, "uniform", 16), (7, 6, "skewed", 8)]


@pytest.mark.parametrize("d,nd,gen,k", ORACLE_CASES)
def test_slab_order_join_equals_oracle(oracle, d, nd, gen, k):
    pts = _points(d, gen, 20_000, seed=d * 10 + nd)
    ids, dist, st = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025)
    rc, i2, d2, s2 = oracle.knnjoin(pts, k, nd=nd, beta=0.025)
    assert rc == 0
    assert np.array_equal(ids, i2)
    assert np.array_equal(bits(dist), bits(d2))


@pytest.mark.parametrize("d,nd,gen", [(4, 2, "uniform"), (4, 2, "skewed"), (6, 4, "uniform"), (5, 3, "lattice"),
                                      (8, 6, "uniform")])
def test_slab_order_off_gives_same_rows(monkeypatch, d, nd, gen):
    pts = _points(d, gen, 200_000, seed=d * 100 + nd + 7)
    loaded = learned_radius.register_default_models()
    if not learned_radius.has_model(loaded, d, nd):
        _lib.set_radius_model(nd, learned_radius.init_weights(d * 10 + nd) * 0.1, 0.3, ndim=d)
    on_p = _lib.knnjoin_ex(pts, 32, num_indexed_dims=nd, beta=0.025)
    on_l = _lib.knnjoin_ex(pts, 32, num_indexed_dims=nd, mode=_lib.KNN_MODE_LEARNED)
    monkeypatch.setenv("KNN_SLAB", "0")
    off_p = _lib.knnjoin_ex(pts, 32, num_indexed_dims=nd, beta=0.025)
    for a in (on_p, on_l):
        assert np.array_equal(a[0], off_p[0])
        assert np.array_equal(bits(a[1]), bits(off_p[1]))
    # the cut streams fewer candidates than whole cells
    assert on_p[2].candidates < off_p[2].candidates


@pytest.mark.parametrize("d,nd", [(4, 2), (6, 4), (8, 6)])
def test_slab_order_double(oracle, d, nd):
    # the slab codes in FP64 arithmetic (the DTYPE=double build of the grid)
    pts = synthetic.skewed_cube(20_000, d, seed=d * 7 + nd).astype(np.float64)
    ids, dist, st = _lib.knnjoin_ex(pts, 8, num_indexed_dims=nd, beta=0.025)
    rc, i2, d2, s2 = oracle.knnjoin(pts, 8, nd=nd, beta=0.025)
    assert rc == 0
    assert np.array_equal(ids, i2)
    assert np.array_equal(dist.view(np.uint64), d2.view(np.uint64))
