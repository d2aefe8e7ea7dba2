"""Python interface with the surface of the reference's knnjoingpu.py (knnjoingpu.py:146-223).

    neighborTable, neighborTableDistances = knnjoin(dataset, KNN, numdim, dtype, verbose)

`dataset` is a flat list/array of N*numdim values (row-major), `dtype` "float" or "double".
Returns two 1-D arrays of N*(KNN+1) entries: int32 neighbour ids and distances, row i =
point i's KNN+1 nearest points (self included) ascending by distance.

Differences from the reference wrapper, all at the interface's edges:
  * the library is found next to this module (the reference loads it from os.getcwd(),
    knnjoingpu.py:173-175);
  * np.asfarray (removed in NumPy 2, knnjoingpu.py:184) is not used;
  * a failed call raises KnnError with the library's status instead of returning zeros.
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import _lib


class HideOutput:
    """Silences fd 1 (the library's stdout) for the scope, as knnjoingpu.py:95-118 does."""

    def __enter__(self):
        sys.stdout.flush()
        self._saved = os.dup(1)
        self._devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(self._devnull, 1)
        return self

    def __exit__(self, *exc):
        sys.stdout.flush()
        os.dup2(self._saved, 1)
        os.close(self._saved)
        os.close(self._devnull)
        return False


def knnjoin(DATASET, KNN, NDIM, DTYPE="float", verbose=False):
    """knnjoingpu.py:146-153"""
    if not verbose:
        with HideOutput():
            return knnjoinmain(DATASET, KNN, NDIM, DTYPE)
    return knnjoinmain(DATASET, KNN, NDIM, DTYPE)


def knnjoinmain(DATASET, KNN, NDIM, DTYPE="float"):
    """knnjoingpu.py:156-223: flat input, ctypes call of KNNJoinPy, flat outputs."""
    if DTYPE not in ("float", "double"):
        raise ValueError("DTYPE must be 'float' or 'double'")
    L = _lib.lib()
    np_t = np.float32 if DTYPE == "float" else np.float64
    data = np.ascontiguousarray(np.asarray(DATASET, dtype=np_t).reshape(-1))
    numPoints = int(len(data) / NDIM)
    print("[Python] Number of rows in file: %d" % (numPoints))
    data = data[: numPoints * NDIM]
    ids = np.zeros(numPoints * (KNN + 1), dtype=np.int32)
    dists = np.zeros(numPoints * (KNN + 1), dtype=np_t)
    fn = L.KNNJoinPy if DTYPE == "float" else L.KNNJoinPyDouble
    fn(_lib.p(data), numPoints, KNN, NDIM, _lib.p(ids), _lib.p(dists))
    _lib.check(L.knnjoin_last_status())
    sumOfDistances = np.sum(dists)
    print("[Python] Sum of distances to all KNN (each distance has been squared using "
          "Euclidean Distance): %f" % (sumOfDistances))
    return ids, dists
