"""GPU tests of the tile blocks of two cell rows (grids of 2 indexed dims): a tile then takes its
queries from two runs of the active list (the block of an even row and the block one row up).
The join equals the one with single-row blocks (KNN_TILE_ROWS=1) and the oracle's, bit for bit,
on uniform, clustered and lattice data, in both schedules; the verbose pass lines show that
the row pairs were chosen.  The cost model's density basis (per cell holding active queries,
or per non-empty cell) changes the tiles and not the rows."""
import ctypes
import re

import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu

_libc = ctypes.CDLL(None)


def _pass_rows(capfd):
    _libc.fflush(None)  # the library prints its pass lines with C stdio
    return [int(m) for m in re.findall(r" rows (\d+) ", capfd.readouterr().out)]


def _points(gen, n, seed):
    if gen == "uniform":
        return synthetic.unit_cube(n, 2, seed=seed)
    if gen == "mixed":
        return synthetic.mixed_box(n, seed=seed)
    if gen == "lattice":
        # coordinates on a 1/512 lattice: equal points and equal keys across a block's two rows
        return (np.round(synthetic.unit_cube(n, 2, seed=seed) * 512.0) / 512.0).astype(np.float32)
    raise ValueError(gen)


# 500k points: the cost model keeps single rows while a pass has too few active queries to fill
# 8192 tiles of a two-row block's size (at 60k uniform points, K=16, it does in every pass)
@pytest.mark.parametrize("gen,k", [("uniform", 16), ("mixed", 32), ("lattice", 8)])
def test_row_pairs_equal_single_rows_and_oracle(oracle, monkeypatch, capfd, gen, k):
    pts = _points(gen, 500_000, seed=k)
    capfd.readouterr()
    ids, dist, _ = _lib.knnjoin_ex(pts, k, beta=0.025, verbose=1)
    assert 2 in _pass_rows(capfd)
    monkeypatch.setenv("KNN_TILE_ROWS", "1")
    ids1, dist1, _ = _lib.knnjoin_ex(pts, k, beta=0.025, verbose=1)
    assert set(_pass_rows(capfd)) == {1}
    assert np.array_equal(ids, ids1)
    assert np.array_equal(dist.view(np.uint32), dist1.view(np.uint32))
    rc, i2, d2, _ = oracle.knnjoin(pts, k, nd=2, beta=0.025)
    assert rc == 0
    assert np.array_equal(ids, i2)
    assert np.array_equal(dist.view(np.uint32), d2.view(np.uint32))


def test_row_pairs_learned_schedule_full_size(monkeypatch, capfd):
    # 2M mixed points, K=64: every learned pass on row pairs or single rows, same rows out
    pts = synthetic.mixed_box(2_000_000, seed=3)
    learned_radius.register_default_models()
    capfd.readouterr()
    ids, dist, _ = _lib.knnjoin_ex(pts, 64, mode=_lib.KNN_MODE_LEARNED, verbose=1)
    assert 2 in _pass_rows(capfd)
    monkeypatch.setenv("KNN_TILE_ROWS", "1")
    ids1, dist1, _ = _lib.knnjoin_ex(pts, 64, mode=_lib.KNN_MODE_LEARNED, verbose=1)
    assert set(_pass_rows(capfd)) == {1}
    assert np.array_equal(ids, ids1)
    assert np.array_equal(dist.view(np.uint32), dist1.view(np.uint32))


@pytest.mark.parametrize("gen,k", [("mixed", 32), ("uniform", 16)])
def test_density_basis_gives_same_rows(monkeypatch, gen, k):
    # the tile-block cost model priced per cell holding active queries (default) or per
    # non-empty cell (KNN_SPAN_DENSITY=cells): other tiles, the same rows
    pts = _points(gen, 500_000, seed=k + 1)
    ids, dist, _ = _lib.knnjoin_ex(pts, k, beta=0.025)
    monkeypatch.setenv("KNN_SPAN_DENSITY", "cells")
    ids1, dist1, _ = _lib.knnjoin_ex(pts, k, beta=0.025)
    assert np.array_equal(ids, ids1)
    assert np.array_equal(dist.view(np.uint32), dist1.view(np.uint32))
