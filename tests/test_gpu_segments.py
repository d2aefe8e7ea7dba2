"""GPU tests of the range kernels' segment streaming (the select kernel's per-query pass 2 and
the collect kernel): a query's rows or cells, 64 per batch, are concatenated and each
candidate finds its segment from a bitmap of the segment starts, 2048 positions per chunk.
KNN_SEG_CHUNK_LOG2 shrinks the chunks (down to 64 positions), so that small inputs cross many
chunk boundaries; the rows must not change, and they equal the oracle's."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32)


# (d, nd, data, k): 3 and 4 indexed dims (select kernel, per-query pass 2, rows), projected
# grids (cells cut to the slab), 6 indexed dims (collect kernel)
CASES = [(3, 3, "uniform", 32), (4, 4, "skewed", 32), (4, 2, "uniform", 16), (6, 4, "skewed", 16),
         (6, 6, "uniform", 16), (8, 6, "skewed", 8)]


def _points(d, gen, n, seed):
    return (synthetic.unit_cube if gen == "uniform" else synthetic.skewed_cube)(n, d, seed=seed)


@pytest.mark.parametrize("d,nd,gen,k", CASES)
def test_small_bitmap_chunks_give_the_same_rows(oracle, monkeypatch, d, nd, gen, k):
    pts = _points(d, gen, 60_000, seed=d * 31 + nd)
    ref = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025)
    for c in (6, 8):
        monkeypatch.setenv("KNN_SEG_CHUNK_LOG2", str(c))
        ids, dist, st = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025)
        assert np.array_equal(ids, ref[0]), c
        assert np.array_equal(bits(dist), bits(ref[1])), c
    monkeypatch.delenv("KNN_SEG_CHUNK_LOG2")
    q = np.random.default_rng(d + nd).choice(len(pts), 300, replace=False)
    perm, _ = oracle.reorder(pts)
    bi, bd = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), k, queries=q)
    assert np.array_equal(ref[0][q], bi)
    assert np.array_equal(bits(ref[1][q]), bits(bd))
