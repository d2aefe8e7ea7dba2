"""The 4-D / 6-D workloads at full size: 5,000,000 float32 points, uniform in [0,1]^d or skewed
(each coordinate Exp(lambda=4) clipped to [0,1]), k=32, NUMINDEXEDDIM = d and d-2, all eight
cases.  Per case: the plain and the learned schedule give the same table bit for bit (with
the shipped radius model for (d, nd), or an untrained one when none ships: rows never depend
on the radius), and 500 seeded random queries equal the oracle's FP64 brute force over all
5M points, ids and distance bits."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu

CASES = [(4, 4, "uniform"), (4, 2, "uniform"), (4, 4, "skewed"), (4, 2, "skewed"),
         (6, 6, "uniform"), (6, 4, "uniform"), (6, 6, "skewed"), (6, 4, "skewed")]


@pytest.mark.parametrize("d,nd,gen", CASES)
def test_highdim_5m_plain_learned_and_bruteforce(oracle, d, nd, gen):
    n, k = 5_000_000, 32
    loaded = learned_radius.register_default_models()
    if not learned_radius.has_model(loaded, d, nd):
        _lib.set_radius_model(nd, learned_radius.init_weights(d * 10 + nd) * 0.1, 0.3, ndim=d)
    make = synthetic.unit_cube if gen == "uniform" else synthetic.skewed_cube
    pts = make(n, d, seed=0x5EED0100 + d * 10 + nd)
    print(f"highdim n={n} d={d} nd={nd} {gen} k={k}", flush=True)
    ip, dp, sp = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, beta=0.025)
    il, dl, sl = _lib.knnjoin_ex(pts, k, num_indexed_dims=nd, mode=_lib.KNN_MODE_LEARNED)
    assert sp.failed_final == 0 and sl.failed_final == 0
    assert np.array_equal(ip, il)
    assert np.array_equal(dp.view(np.uint32), dl.view(np.uint32))
    assert sp.checksum == sl.checksum
    del il, dl
    q = np.random.default_rng(d * nd + (7 if gen == "skewed" else 0)).choice(n, 500, replace=False)
    perm, _ = oracle.reorder(pts)
    assert list(perm) == list(sp.perm[:d])
    bi, bd = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), k, queries=q)
    assert np.array_equal(ip[q], bi)
    assert np.array_equal(dp[q].view(np.uint32), bd.view(np.uint32))
