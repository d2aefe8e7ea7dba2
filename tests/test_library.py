"""CPU tests of the C-ABI library: it loads and exports every declared symbol (no GPU calls)."""
import ctypes
import os
import re

from hybrid_k_nearest_neighbor_self_join_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpuknnselfjoin.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    return sorted({n for n in names if n.startswith(("knnjoin_", "KNNJoin"))})


def test_header_symbols_listed():
    assert declared_symbols() == sorted(_lib.EXPORTED_SYMBOLS)


def test_library_loads_and_exports_everything():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_default_params_match_reference_defaults():
    prm = _lib.default_params()
    assert prm.beta == 0.05 and prm.reorder == 1 and prm.num_indexed_dims == 0
    assert prm.mode == _lib.KNN_MODE_BETA


def test_double_build_exports_everything_with_a_double_knnjoinpy():
    # the DTYPE=double build (params.h:5): same symbols, KNNJoinPy takes double*
    assert os.path.exists(_lib.LIB_DOUBLE_PATH)
    L = ctypes.CDLL(_lib.LIB_DOUBLE_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name
    L.knnjoin_py_dtype.restype = ctypes.c_int
    assert L.knnjoin_py_dtype() == _lib.KNN_FLOAT64
    assert _lib.lib().knnjoin_py_dtype() == _lib.KNN_FLOAT32
