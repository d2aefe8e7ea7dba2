"""knnjoin_ex_noalloc (no device allocation; buffers in the library's static scratch array)
against knnjoin_ex on the same device-resident inputs: rows and distance bits equal, on two
calls in a row, the second with a smaller input, for both radius schedules; inputs beyond
the scratch's limit are rejected with a clear error before any work.

Device buffers come from hipMalloc of the HIP runtime the library itself is linked against
(the process has loaded it with the library), so the test needs no other GPU framework."""
import ctypes

import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.lib()  # the library's HIP runtime is in the process from here on
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h, self.p = h, ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def free(self):
        self.h.hipFree(self.p)


def _join(fn, pts, k, mode):
    h = _hip()
    n, d = pts.shape
    dp, di, dd = DevBuf(h, pts.nbytes), DevBuf(h, n * (k + 1) * 4), DevBuf(h, n * (k + 1) * 4)
    try:
        assert h.hipMemcpy(dp.p, pts.ctypes.data, pts.nbytes, H2D) == 0
        assert h.hipMemset(di.p, 0xF9, di.nbytes) == 0  # ids -7 where nothing is written
        assert h.hipMemset(dd.p, 0, dd.nbytes) == 0
        prm = _lib.default_params(beta=0.025, mode=mode, device_ptrs=1)
        st = _lib.KnnStats()
        rc = getattr(_lib.lib(), fn)(dp.p, _lib.KNN_FLOAT32, n, d, k, ctypes.byref(prm), di.p, dd.p,
                                     ctypes.byref(st))
        assert h.hipDeviceSynchronize() == 0
        ids = np.empty((n, k + 1), np.int32)
        dists = np.empty((n, k + 1), np.float32)
        assert h.hipMemcpy(ids.ctypes.data, di.p, ids.nbytes, D2H) == 0
        assert h.hipMemcpy(dists.ctypes.data, dd.p, dists.nbytes, D2H) == 0
        return rc, ids, dists, st
    finally:
        for b in (dp, di, dd):
            b.free()


@pytest.mark.parametrize("mode", [_lib.KNN_MODE_BETA, _lib.KNN_MODE_LEARNED])
def test_noalloc_matches_knnjoin_ex_twice(mode):
    if mode == _lib.KNN_MODE_LEARNED and not learned_radius.has_model(learned_radius.register_default_models(), 2, 2):
        pytest.skip("no 2-D radius model")
    for n, k, seed in [(400_000, 32, 31), (120_000, 16, 32)]:
        pts = synthetic.mixed_box(n, seed=seed)
        print(f"noalloc n={n} k={k} mode={mode}", flush=True)
        rc0, i0, d0, s0 = _join("knnjoin_ex", pts, k, mode)
        rc1, i1, d1, s1 = _join("knnjoin_ex_noalloc", pts, k, mode)
        assert rc0 == 0 and rc1 == 0, _lib.lib().knnjoin_last_error()
        assert np.array_equal(i0, i1)
        assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
        assert s0.checksum == s1.checksum and s1.failed_final == 0


def test_noalloc_rejects_inputs_beyond_its_scratch():
    pts = synthetic.uniform_box(3_000_000, seed=33)
    rc, ids, _, _ = _join("knnjoin_ex_noalloc", pts, 8, _lib.KNN_MODE_BETA)
    assert rc == _lib.KNN_ERR_NOMEM
    assert b"scratch" in _lib.lib().knnjoin_last_error()
    assert np.all(ids == -0x06060607)  # nothing ran (bytes 0xF9)
    small = synthetic.uniform_box(5000, seed=34)
    rc, _, _, _ = _join("knnjoin_ex_noalloc", small, 100, _lib.KNN_MODE_BETA)
    assert rc == _lib.KNN_ERR_ARG
    # host pointers are refused
    prm = _lib.default_params(device_ptrs=0)
    ids = np.zeros((5000, 9), np.int32)
    d = np.zeros((5000, 9), np.float32)
    rc = _lib.lib().knnjoin_ex_noalloc(_lib.p(small), _lib.KNN_FLOAT32, 5000, 2, 8, ctypes.byref(prm),
                                       _lib.p(ids), _lib.p(d), None)
    assert rc == _lib.KNN_ERR_ARG
