"""Exact K-nearest-neighbour self-join on one AMD MI355X (gfx950), hand-written HIP.

Entry points mirror mgowanlock/hybrid_k_nearest_neighbor_self_join:
  * C ABI `KNNJoinPy` / `knnjoin_ex` in libgpuknnselfjoin.so (include/gpuknnselfjoin.h);
  * `knnjoingpu.knnjoin(dataset, KNN, numdim, dtype, verbose)` (the reference's Python API);
  * the `knnjoin_main <FNAME> <DIM> <K>` CLI.
"""
from . import _lib  # noqa: F401
from .knnjoingpu import knnjoin  # noqa: F401

__all__ = ["knnjoin"]
