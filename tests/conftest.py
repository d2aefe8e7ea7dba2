import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def golden():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_small.npz")
    z = np.load(path, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in z.files})
    out = {}
    for nm in names:
        k, nd = (int(v) for v in z[f"{nm}__k"])
        out[nm] = dict(points=z[f"{nm}__points"], k=k, nd=nd, perm=z[f"{nm}__perm"],
                       ids=z[f"{nm}__ids"], dists=z[f"{nm}__dists"])
    return out


@pytest.fixture(scope="session")
def gpulib():
    from hybrid_k_nearest_neighbor_self_join_amd import _lib
    return _lib
