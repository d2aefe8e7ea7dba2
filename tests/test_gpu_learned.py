"""GPU tests of the learned per-query radius: MLP forward / gradient / training against the
CPU oracle (bitwise: same fmaf order), features (tolerance: log1pf of the GPU's and glibc's
math libraries may differ by an ulp), exactness of the learned schedule whatever the model
predicts, and what the trained model gains over the plain schedule."""
import numpy as np
import pytest

from hybrid_k_nearest_neighbor_self_join_amd import _lib, learned_radius, synthetic

pytestmark = pytest.mark.gpu


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def test_mlp_forward_matches_oracle(oracle):
    rng = np.random.default_rng(1)
    W = learned_radius.init_weights(3) + rng.normal(0, 0.01, _lib.MLP_NW).astype(np.float32)
    X = rng.normal(0, 2, (3000, 12)).astype(np.float32)
    out = _lib.mlp_forward(W, X)
    assert np.array_equal(bits(out), bits(oracle.mlp_forward(W, X)))


@pytest.mark.parametrize("B", [1, 100, 4096])
def test_mlp_gradient_matches_oracle(oracle, B):
    rng = np.random.default_rng(B)
    W = learned_radius.init_weights(B)
    W[768:832] = rng.normal(0, 0.1, 64)
    X = rng.normal(0, 1, (B, 12)).astype(np.float32)
    Y = rng.normal(0, 1, B).astype(np.float32)
    g, loss = _lib.mlp_grad(W, X, Y)
    g2, loss2 = oracle.mlp_grad(W, X, Y)
    assert np.array_equal(bits(g), bits(g2))
    assert loss == loss2


def test_mlp_training_matches_oracle(oracle):
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (600, 12)).astype(np.float32)
    Y = (X[:, 0] * 0.5 - X[:, 3] + 0.1).astype(np.float32)
    W0 = learned_radius.init_weights(11)
    w, losses = _lib.mlp_train(W0, X, Y, epochs=3, batch=128, lr=1e-3, seed=5)
    w2 = oracle.mlp_train(W0, X, Y, 3, 128, 1e-3, 5)
    assert np.array_equal(bits(w), bits(w2))
    assert losses[-1] < losses[0]


@pytest.mark.parametrize("gen,nd,k,reps", [("mixed", 2, 64, None), ("mixed", 2, 64, "1"), ("mixed", 2, 64, "3"),
                                            ("uniform", 2, 8, None), ("cube4", 4, 32, None), ("skew6", 4, 16, None)])
def test_features_match_oracle(oracle, monkeypatch, gen, nd, k, reps):
    # reps: the point counts over that many replica grids (KNN_FEATURE_REPS; None = the
    # library's choice, 16 at these sizes), summed before any feature reads them
    if reps is not None:
        monkeypatch.setenv("KNN_FEATURE_REPS", reps)
    pts = {"mixed": lambda: synthetic.mixed_box(200000, seed=4),
           "uniform": lambda: synthetic.uniform_box(100000, seed=5),
           "cube4": lambda: synthetic.unit_cube(60000, 4, seed=6),
           "skew6": lambda: synthetic.skewed_cube(60000, 6, seed=7)}[gen]()
    f, wf = _lib.radius_features(pts, k, nd=nd)
    perm, _ = oracle.reorder(pts)
    f2, wf2 = oracle.radius_features(pts[:, perm], k, nd)
    assert wf == wf2
    # counts are exact integers: their log1p agrees to float rounding
    assert np.allclose(f, f2, rtol=2e-7, atol=0)
    assert np.array_equal(np.round(np.expm1(f[:, :10])), np.round(np.expm1(f2[:, :10])))


@pytest.mark.parametrize("scale", [0.5, 2.0])
def test_radius_inference_matches_oracle_mlp(oracle, scale):
    # the join's radius inference (MLP on the matrix cores, FP16 operands, FP32 accumulation)
    # against the oracle's FP32 MLP on the same features: log-radius within 3e-2, 5e-3 at the
    # median; the radius only chooses which pass a query is first tried in
    pts = synthetic.mixed_box(50000, seed=12)
    w = learned_radius.init_weights(17) * scale
    margin = 0.25
    _lib.set_radius_model(2, w, margin)
    f, wf, r = _lib.radius_features(pts, 16, nd=2, want_radius=True)
    y = oracle.mlp_forward(w, f)
    e = np.clip(y.astype(np.float64) + margin, -8.0, 14.0)
    err = np.abs(np.log(r.astype(np.float64) / (wf * np.exp(e))))
    assert err.max() < 3e-2 and np.median(err) < 5e-3, (err.max(), np.median(err))
    # spread of the prediction itself, so that the bound above is not vacuous
    assert np.std(y) > 0.05


def _random_model(nd, seed, scale):
    w = learned_radius.init_weights(seed) * scale
    _lib.set_radius_model(nd, w, 0.0)


@pytest.mark.parametrize("scale", [0.0, 0.3, 3.0])
def test_learned_schedule_exact_with_any_model(oracle, scale):
    # whatever the network predicts (constant, noise, wild), rows equal the oracle's
    pts = synthetic.mixed_box(120000, seed=21)
    _random_model(2, 99, scale)
    # brute_fallback=0 keeps every query on the learned passes (the fallback is tested apart)
    ids, d, st = _lib.knnjoin_ex(pts, 32, mode=_lib.KNN_MODE_LEARNED, brute_fallback=0)
    assert st.brute_queries == 0
    rc, i2, d2, _ = oracle.knnjoin(pts, 32, nd=2, beta=0.025)
    assert rc == 0
    assert np.array_equal(ids, i2) and np.array_equal(bits(d), bits(d2))


def test_learned_schedule_exact_higher_dims(oracle):
    pts = synthetic.skewed_cube(50000, 4, seed=3)
    _random_model(4, 5, 0.5)
    ids, d, st = _lib.knnjoin_ex(pts, 16, num_indexed_dims=4, mode=_lib.KNN_MODE_LEARNED,
                                 brute_fallback=0)
    rc, i2, d2, _ = oracle.knnjoin(pts, 16, nd=4, beta=0.025)
    assert np.array_equal(ids, i2) and np.array_equal(bits(d), bits(d2))


def test_missing_model_is_an_error():
    pts = synthetic.unit_cube(5000, 7, seed=1)
    with pytest.raises(_lib.KnnError) as e:
        _lib.knnjoin_ex(pts, 4, num_indexed_dims=7, mode=_lib.KNN_MODE_LEARNED)
    assert e.value.code == _lib.KNN_ERR_MODEL


def test_trained_model_gains_over_plain_schedule(oracle):
    loaded = learned_radius.register_default_models()
    if not learned_radius.has_model(loaded, 2, 2):
        pytest.skip("no trained 2-D model file in the package")
    pts = synthetic.mixed_box(2_000_000, seed=0xFEED)
    ids_p, d_p, st_p = _lib.knnjoin_ex(pts, 64, beta=0.025)
    ids_l, d_l, st_l = _lib.knnjoin_ex(pts, 64, mode=_lib.KNN_MODE_LEARNED)
    assert np.array_equal(ids_p, ids_l) and np.array_equal(bits(d_p), bits(d_l))
    sub = np.random.default_rng(0).choice(len(pts), 2000, replace=False)
    bi, _ = oracle.bruteforce(pts, 64, queries=sub)
    assert np.array_equal(ids_l[sub], bi)
    # the gain: fewer passes and fewer (query, candidate) evaluations than the BETA schedule
    assert st_l.iterations < st_p.iterations
    assert st_l.candidates < st_p.candidates
