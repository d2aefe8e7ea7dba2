"""CPU tests: the C oracle pinned against golden fixtures, an independent numpy
restatement, closed forms and invariants (no GPU needed)."""
import math

import numpy as np
import pytest

import np_ref

F32 = np.float32


# ---------------------------------------------------------------- golden fixtures
def test_golden_numpy_restatement_reproduces_fixtures(golden):
    # the fixture generator's own restatement is deterministic: re-deriving one case agrees
    g = golden["lattice2d_k12"]
    perm, _ = np_ref.reorder_perm(g["points"])
    ids, d = np_ref.bruteforce(np.ascontiguousarray(g["points"][:, perm]), g["k"])
    assert np.array_equal(ids, g["ids"]) and np.array_equal(d.view(np.uint32), g["dists"].view(np.uint32))


@pytest.mark.parametrize("beta", [0.0, 0.025, 0.05, 0.2])
def test_oracle_join_matches_golden(oracle, golden, beta):
    for name, g in golden.items():
        rc, ids, dists, st = oracle.knnjoin(g["points"], g["k"], nd=g["nd"], beta=beta, nthreads=4)
        assert rc == 0, name
        assert list(st.perm[: g["points"].shape[1]]) == list(g["perm"]), name
        assert np.array_equal(ids, g["ids"]), name
        assert np.array_equal(dists.view(np.uint32), g["dists"].view(np.uint32)), name


def test_oracle_bruteforce_matches_golden(oracle, golden):
    for name, g in golden.items():
        pr = np.ascontiguousarray(g["points"][:, g["perm"]])
        ids, dists = oracle.bruteforce(pr, g["k"], nthreads=4)
        assert np.array_equal(ids, g["ids"]), name
        assert np.array_equal(dists.view(np.uint32), g["dists"].view(np.uint32)), name


# ---------------------------------------------------------------- reorder (main.cu:2521-2615)
def test_reorder_matches_numpy_and_known_variances(oracle):
    rng = np.random.default_rng(3)
    pts = (rng.uniform(0, 1, (200000, 4)) * np.array([1.0, 4.0, 2.0, 0.5])).astype(np.float32)
    perm, var = oracle.reorder(pts)
    p2, v2 = np_ref.reorder_perm(pts)
    assert list(perm) == list(p2)
    assert np.array_equal(var.view(np.uint32), v2.view(np.uint32))
    # closed form: var of U[0,a] is a^2/12 -> order by a descending
    assert list(perm) == [1, 2, 0, 3]
    assert np.allclose(var, np.array([16, 4, 1, 0.25]) / 12.0, rtol=0.1)


def test_reorder_is_stable_on_ties(oracle):
    pts = np.zeros((1000, 3), np.float32)
    perm, _ = oracle.reorder(pts)
    assert list(perm) == [0, 1, 2]
    pts[:, 2] = np.arange(1000) % 7
    perm, _ = oracle.reorder(pts)
    assert list(perm) == [2, 0, 1]


# ---------------------------------------------------------------- epsilon estimate
def test_avg_dist_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 10, (60000, 2)).astype(np.float32)
    a, c = oracle.avg_dist(pts)
    a2, c2 = np_ref.avg_dist(pts)
    assert c == c2 == sum(len(range(g % 256, 60000, 1000)) for g in range(256) if g * 1000 < 60000)
    assert a == a2


def test_avg_dist_closed_form_unit_square(oracle):
    # mean distance of two uniform points in the unit square = 0.5214054...
    pts = np.random.default_rng(9).uniform(0, 1, (400000, 2)).astype(np.float32)
    a, _ = oracle.avg_dist(pts)
    assert abs(a - 0.52140543) < 0.01


def test_estimate_eps_matches_numpy(oracle):
    rng = np.random.default_rng(7)
    pts = rng.uniform(0, 100, (20000, 2)).astype(np.float32)
    for k, beta in [(8, 0.025), (16, 0.0), (4, 0.2)]:
        rc, info, hist = oracle.estimate_eps(pts, k, beta)
        eps, bucket, h2, avg = np_ref.kdist_eps(pts, k, beta)
        assert rc == 0
        assert info.avg == avg
        assert np.array_equal(hist, h2)
        assert info.bucket == bucket
        assert info.eps == eps
        assert info.offset == 100 and info.nsample == 200


def test_estimate_eps_thresholds():
    # point_threshold_beta = k + (100k - k)*beta truncated (GPU.cu:3499-3501)
    for k, beta, thr in [(8, 0.025, 27), (16, 0.025, 55), (64, 0.025, 222), (128, 0.025, 444), (32, 0.025, 111)]:
        assert int(k + (100 * k - k) * beta) == thr


def test_estimate_eps_density_closed_form(oracle):
    # uniform density rho in 2-D: about (thr+?) neighbours inside pi*eps^2*rho
    n = 200000
    pts = np.random.default_rng(11).uniform(0, 100, (n, 2)).astype(np.float32)
    rc, info, _ = oracle.estimate_eps(pts, 8, 0.025)
    rho = n / 1e4
    # cumulative per sampled point ~= pi r^2 rho (the x8 weight undoes the 1/8 sample)
    r_expected = math.sqrt(info.threshold / (math.pi * rho))
    assert abs(info.eps - r_expected) / r_expected < 0.05


def test_estimate_eps_all_identical_is_zero(oracle):
    pts = np.ones((1000, 2), np.float32)
    rc, info, _ = oracle.estimate_eps(pts, 4, 0.05)
    assert rc == oracle.ORC_ERR_EPS_ZERO and info.eps == 0.0


# ---------------------------------------------------------------- grid + index
def test_grid_dims_and_cell_coord(oracle):
    mn = np.array([0.5, -90.0], np.float32)
    mx = np.array([359.9, 90.0], np.float32)
    eps = 2.3616
    a, b, tot = oracle.grid_dims(mn, mx, eps)
    a2, b2, tot2 = np_ref.grid_dims(mn, mx, eps)
    assert np.array_equal(a.view(np.uint32), a2.view(np.uint32))
    assert list(b) == list(b2) and tot == tot2
    for x in [0.5, 1.0, 100.25, 359.9]:
        assert oracle.cell_coord(x, float(a[0]), eps, int(b[0])) == np_ref.cell_coord(x, a[0], eps, int(b[0]))
    assert oracle.cell_coord(-1e9, 0.0, 1.0, 10) == 0
    assert oracle.cell_coord(1e9, 0.0, 1.0, 10) == 9


@pytest.mark.parametrize("d,nd", [(2, 2), (3, 3), (4, 2), (6, 6)])
def test_index_matches_numpy_and_invariants(oracle, d, nd):
    rng = np.random.default_rng(d * 10 + nd)
    pts = rng.uniform(0, 1, (5000, d)).astype(np.float32)
    eps = 0.11
    mn, mx = oracle.minmax(pts, nd)
    assert np.array_equal(mn, pts[:, :nd].min(0)) and np.array_equal(mx, pts[:, :nd].max(0))
    minArr, nCells, tot = oracle.grid_dims(mn, mx, eps)
    key, sid, ckeys, cstart = oracle.build_index(pts, nd, eps, minArr, nCells)
    k2, sid2, ck2, cs2 = np_ref.build_index(pts, nd, eps, minArr, nCells)
    assert np.array_equal(key, k2) and np.array_equal(sid, sid2)
    assert np.array_equal(ckeys, ck2) and np.array_equal(cstart, cs2)
    # invariants: permutation; cell table sorted & unique; ids ascending inside a cell
    assert np.array_equal(np.sort(sid), np.arange(5000))
    assert np.all(np.diff(ckeys.astype(np.int64)) > 0) and ckeys.max() < tot
    for c in range(len(ckeys)):
        seg = sid[cstart[c]:cstart[c + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        assert np.all(key[seg] == ckeys[c])


# ---------------------------------------------------------------- whole plain method
@pytest.mark.parametrize("gen,k,nd", [
    ("uniform2d", 8, 2), ("mixed2d", 32, 2), ("cube3d", 5, 3), ("cube6d", 16, 4), ("skew4d", 32, 2),
])
def test_oracle_join_equals_bruteforce(oracle, gen, k, nd):
    rng = np.random.default_rng(hash(gen) % 2**32)
    if gen == "uniform2d":
        pts = np.column_stack([rng.uniform(0, 360, 6000), rng.uniform(-90, 90, 6000)])
    elif gen == "mixed2d":
        u = rng.uniform(0, 360, (4000, 2))
        c = rng.uniform(0, 360, (16, 2))[rng.integers(0, 16, 2000)] + rng.normal(0, 0.5, (2000, 2))
        pts = np.concatenate([u, c])
    elif gen == "cube3d":
        pts = rng.uniform(0, 1, (4000, 3))
    elif gen == "cube6d":
        pts = rng.uniform(0, 1, (3000, 6))
    else:
        pts = np.clip(rng.exponential(0.25, (4000, 4)), 0, 1)
    pts = pts.astype(np.float32)
    rc, ids, dists, st = oracle.knnjoin(pts, k, nd=nd, beta=0.025, nthreads=8)
    assert rc == 0
    perm = np.array(st.perm[: pts.shape[1]])
    bi, bd = oracle.bruteforce(np.ascontiguousarray(pts[:, perm]), k, nthreads=8)
    assert np.array_equal(ids, bi)
    assert np.array_equal(dists.view(np.uint32), bd.view(np.uint32))
    # invariants: self included, rows ascending by distance up to FP32 rounding
    assert np.all((ids == np.arange(len(pts))[:, None]).any(1))
    assert np.all(np.diff(dists.astype(np.float64), axis=1) >= -1e-6 * (1 + dists[:, 1:]))


def test_oracle_results_independent_of_beta(oracle):
    pts = np.random.default_rng(2).uniform(0, 10, (5000, 2)).astype(np.float32)
    ref = None
    iters = []
    for beta in [0.0, 0.025, 0.2, 1.0]:
        rc, ids, d, st = oracle.knnjoin(pts, 12, beta=beta, nthreads=8)
        assert rc == 0
        iters.append(st.iterations)
        if ref is None:
            ref = (ids, d)
        assert np.array_equal(ids, ref[0]) and np.array_equal(d, ref[1])
    assert iters[0] >= iters[-1]


def test_oracle_error_paths(oracle):
    rc, *_ = oracle.knnjoin(np.zeros((5, 2), np.float32) + np.arange(5, dtype=np.float32)[:, None], 5)
    assert rc == oracle.ORC_ERR_TOO_FEW


def test_oracle_degenerate_eps_fallback(oracle):
    # all points identical: every distance is 0, rows are the K+1 smallest ids
    rc, ids, d, st = oracle.knnjoin(np.ones((100, 2), np.float32), 3)
    assert rc == 0 and st.eps0 == 1.0
    assert np.array_equal(ids, np.tile(np.arange(4, dtype=np.int32), (100, 1)))
    assert np.all(d == 0)
    # N <= 1000: the reference's distance sample is point 0 alone -> BETA estimate 0
    pts = np.random.default_rng(0).uniform(0, 1, (500, 2)).astype(np.float32)
    rc, info, _ = oracle.estimate_eps(pts, 4, 0.05)
    assert rc == oracle.ORC_ERR_EPS_ZERO and info.avg == 0.0
    rc, ids, d, st = oracle.knnjoin(pts, 4)
    assert rc == 0 and st.eps0 > 0
    ext = max(F32(pts[:, 0].max() - pts[:, 0].min()), F32(pts[:, 1].max() - pts[:, 1].min()))
    assert st.eps0 == F32(float(ext) * (5.0 / 500.0) ** 0.5)


def test_grow_rule(oracle):
    for eps, eps0 in [(0.125, 0.125), (1e-3, 7e-4), (3.3, 1.1)]:
        want = F32(np.float64(F32(eps)) + np.float64(F32(eps0)) * 0.5)
        assert oracle.grow_eps(eps, eps0) == want


def test_checksum_definition(oracle):
    d = np.random.default_rng(1).uniform(0, 1, (5000, 9)).astype(np.float32)
    want, chunk = 0.0, 0.0
    for r in range(d.shape[0]):
        rs = 0.0
        for v in d[r]:
            rs += float(F32(v * v))
        chunk += rs
        if (r + 1) % 4096 == 0 or r + 1 == d.shape[0]:
            want += chunk
            chunk = 0.0
    assert oracle.checksum(d) == want
    rc, ids, dd, st = oracle.knnjoin(np.random.default_rng(4).uniform(0, 1, (9000, 2)).astype(np.float32), 6)
    assert st.checksum == oracle.checksum(dd)


# ---------------------------------------------------------------- learned-radius MLP oracle
def _np_mlp(W, X):
    W = W.astype(np.float64)
    W1 = W[:768].reshape(64, 12)
    b1, W2, b2, w3, b3 = W[768:832], W[832:4928].reshape(64, 64), W[4928:4992], W[4992:5056], W[5056]
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ w3 + b3


def test_oracle_mlp_forward_and_gradient_closed_form(oracle):
    rng = np.random.default_rng(0)
    W = rng.normal(0, 0.2, 5057).astype(np.float32)
    X = rng.normal(0, 1, (200, 12)).astype(np.float32)
    Y = rng.normal(0, 1, 200).astype(np.float32)
    out = oracle.mlp_forward(W, X)
    assert np.allclose(out, _np_mlp(W, X.astype(np.float64)), rtol=1e-4, atol=1e-5)
    G, loss = oracle.mlp_grad(W, X, Y)

    def L(Wv):
        return np.mean((_np_mlp(Wv, X.astype(np.float64)) - Y) ** 2)

    assert abs(loss - L(W)) < 1e-5 * max(1.0, L(W))
    Wd = W.astype(np.float64)
    for p in [0, 11, 700, 770, 831, 900, 3000, 4927, 4930, 4991, 5000, 5056]:
        e = 1e-4
        a, b = Wd.copy(), Wd.copy()
        a[p] += e
        b[p] -= e
        fd = (L(a) - L(b)) / (2 * e)
        assert abs(G[p] - fd) <= 1e-4 * max(1.0, abs(fd)), (p, G[p], fd)


def test_oracle_adam_first_two_steps_closed_form(oracle):
    # Adam (Kingma & Ba 2015) with beta1 0.9, beta2 0.999, eps 1e-8 and bias correction: at
    # step 1 every weight moves by exactly -lr g1 / (|g1| + eps) (m^ = g1, v^ = g1^2); at step 2
    # by -lr m^2 / (sqrt(v^2) + eps) with m2 = 0.9 (0.1 g1) + 0.1 g2, v2 = 0.999 (0.001 g1^2) +
    # 0.001 g2^2, m^2 = m2 / (1 - 0.9^2), v^2 = v2 / (1 - 0.999^2).  A dropped bias correction
    # scales step 1 by 0.1 / sqrt(0.001); swapped betas change step 2.  One batch of all the
    # samples per epoch, so the epoch shuffle only reorders the batch's terms (rounding).
    rng = np.random.default_rng(7)
    W0 = rng.normal(0, 0.2, 5057).astype(np.float32)
    X = rng.normal(0, 1, (64, 12)).astype(np.float32)
    Y = rng.normal(0, 1, 64).astype(np.float32)
    lr = 1e-3
    W1 = oracle.mlp_train(W0, X, Y, 1, 64, lr, 11)
    W2 = oracle.mlp_train(W0, X, Y, 2, 64, lr, 11)
    g1, _ = oracle.mlp_grad(W0, X, Y)
    g2, _ = oracle.mlp_grad(W1, X, Y)
    g1, g2 = g1.astype(np.float64), g2.astype(np.float64)
    eps = 1e-8
    step1 = -lr * g1 / (np.abs(g1) + eps)
    m2 = 0.9 * (0.1 * g1) + 0.1 * g2
    v2 = 0.999 * (0.001 * g1 * g1) + 0.001 * g2 * g2
    step2 = -lr * (m2 / (1 - 0.9 ** 2)) / (np.sqrt(v2 / (1 - 0.999 ** 2)) + eps)
    big1 = np.abs(g1) > 1e-4 * np.abs(g1).max()
    big2 = np.abs(g2) > 1e-4 * np.abs(g2).max()
    assert big1.sum() > 3000 and big2.sum() > 3000
    d1 = W1.astype(np.float64) - W0
    d2 = W2.astype(np.float64) - W1
    # float32 weights: a step of lr is resolved to about 2^-24 |W| / lr of itself
    tol = 2e-3 * lr + 4 * 2.0 ** -24 * np.abs(W0.astype(np.float64))
    assert np.all(np.abs(d1 - step1)[big1] <= tol[big1])
    assert np.all(np.abs(d1) <= lr * (1 + 1e-3) + tol)
    assert np.all(np.abs(d2 - step2)[big1 & big2] <= (tol + 2e-3 * np.abs(step2))[big1 & big2])


def test_oracle_features_definition(oracle):
    # lattice with known counts: 3 points per unit cell in a 10x10 box
    g = np.stack(np.meshgrid(np.arange(10), np.arange(10)), -1).reshape(-1, 2).astype(np.float32)
    pts = np.concatenate([g + 0.25, g + 0.5, g + 0.75]).astype(np.float32)
    f, wf = oracle.radius_features(pts, 2, 2)
    assert f.shape == (300, 12)
    lo = pts.min(0) - np.float32(wf)
    cells = np.floor((pts - lo) / np.float32(wf)).astype(np.int64)
    occ = len(np.unique(cells, axis=0))
    assert np.allclose(f[:, 10], np.log(3.0)) and np.allclose(f[:, 11], np.log(300 / occ))
    # own-cell counts are the lattice's: every point's cell holds the points sharing it
    _, inv, cnt = np.unique(cells, axis=0, return_inverse=True, return_counts=True)
    assert np.allclose(f[:, 0], np.log1p(cnt[inv.ravel()]))
    assert np.all(f[:, 0] > 0) and np.all(f[:, 1:10].max(1) >= f[:, 0] - 1e-6)
