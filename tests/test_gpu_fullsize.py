"""The headline workload at full size: 10,000,000 mixed-density 2-D float32 points (70 % uniform
over [0,360)x[-90,90], 30 % in 64 Gaussian clusters), k = 16, 64, 128, BETA=0.025.

Checks per k:
  * the learned and the plain schedule give the same table bit for bit (results do not depend
    on the radius schedule, main.cu:305-430);
  * 100,000 seeded random queries equal the oracle's plain-method grid join restricted to
    those queries (the same reorder, BETA epsilon, grids over all 10M points and
    failed-query loop, main.cu:109-510), ids and distance bits;
  * 1,000 seeded random queries equal the oracle's FP64 brute force over all 10M points;
  * over the whole table: every row holds the query itself, ids are in range, FP32 distances
    are non-decreasing up to their rounding (two neighbours whose FP64 keys differ by less
    than a few FP32 ulps may swap their FP32 values), and the checksum equals the oracle's
    checksum of the table (main.cu:452-455)."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [16, 64, 128])
def test_full_size_tables_agree_with_oracle(oracle, k):
    if not learned_radius.has_model(learned_radius.register_default_models(), 2, 2):
        pytest.skip("no 2-D radius model")
    n = 10_000_000
    pts = synthetic.mixed_box(n, seed=0x5EED0002)
    print(f"full-size join n={n} k={k}", flush=True)
    il, dl, sl = _lib.knnjoin_ex(pts, k, mode=_lib.KNN_MODE_LEARNED)
    ip, dp, sp = _lib.knnjoin_ex(pts, k, beta=0.025)
    assert sl.failed_final == 0 and sp.failed_final == 0
    assert np.array_equal(il, ip)
    assert np.array_equal(dl.view(np.uint32), dp.view(np.uint32))
    assert sl.checksum == sp.checksum
    del ip, dp

    assert np.all(dl[:, 1:] >= dl[:, :-1] * np.float32(1.0 - 2.0**-21))
    assert np.all((il == np.arange(n, dtype=np.int32)[:, None]).any(axis=1))
    assert il.min() >= 0 and il.max() < n
    assert oracle.checksum(dl) == sl.checksum

    # 100k seeded queries: the oracle's grid join of exactly those queries
    rng = np.random.default_rng(0x5EED0003 + k)
    sub = np.sort(rng.choice(n, 100_000, replace=False)).astype(np.uint32)
    print(f"oracle subset join: {len(sub)} queries", flush=True)
    rc, oi, od, ost = oracle.knnjoin(pts, k, nd=2, beta=0.025, subset=sub)
    assert rc == 0
    assert list(ost.perm[:2]) == list(sl.perm[:2])
    bad = np.flatnonzero((il[sub] != oi).any(axis=1))
    assert bad.size == 0, f"{bad.size} of {len(sub)} rows differ, first query {sub[bad[0]]}"
    assert np.array_equal(dl[sub].view(np.uint32), od.view(np.uint32))
    del oi, od

    # 1000 seeded queries: FP64 brute force over every point
    perm = list(sl.perm[:2])
    q = rng.choice(n, 1000, replace=False).astype(np.uint32)
    bi, bd = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), k, queries=q)
    assert np.array_equal(il[q], bi)
    assert np.array_equal(dl[q].view(np.uint32), bd.view(np.uint32))
