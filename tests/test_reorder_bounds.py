"""The join decides the dimension order from FP64 moments when the error intervals of the
sequential FP32 variances (main.cu:2537-2562) are pairwise disjoint, and falls back to the
sequential sums otherwise.  This checks the interval rule (restated here in numpy) against the
oracle's exact sequential FP32 variances on random data: every exact variance lies inside its
interval, and every order the rule decides equals the oracle's."""
import numpy as np


def intervals(pts):
    n, d = pts.shape
    x = pts[::100].astype(np.float64)
    m = len(x)
    u = 2.0 ** -24
    g = (m + 4) * u * 1.01
    den = n / 100.0
    S, A = x.sum(0), np.abs(x).sum(0)
    a = S / den
    Q, C = ((x - a) ** 2).sum(0), (x - a).sum(0)
    da = (g + 4 * u) * A / den
    slack = 1e-9 * (Q + A * A / m)
    lo = (Q - 2 * da * np.abs(C) - slack) * (1 - g) * (1 - 4 * u) / den
    hi = (Q + 2 * da * np.abs(C) + m * da * da + slack) * (1 + g) * (1 + 4 * u) / den
    return lo, hi, Q / den


def test_interval_rule_against_exact_float_variances(oracle):
    rng = np.random.default_rng(20261202)
    decided = 0
    for _ in range(150):
        n = int(rng.integers(1000, 200000))
        d = int(rng.integers(2, 6))
        scale = rng.uniform(0.5, 2.0, d) ** rng.integers(1, 8)
        off = rng.normal(0, 1, d) * rng.choice([0.0, 1.0, 1e3, 1e5])
        pts = (rng.normal(0, 1, (n, d)) * scale + off).astype(np.float32)
        perm, var = oracle.reorder(pts)
        v = np.empty(d)
        v[perm] = var.astype(np.float64)  # oracle returns the variances in sorted order
        lo, hi, mid = intervals(pts)
        assert np.all((v >= lo) & (v <= hi))
        if all(lo[i] > hi[j] or lo[j] > hi[i] for i in range(d) for j in range(i + 1, d)):
            decided += 1
            assert sorted(range(d), key=lambda i: -mid[i]) == list(perm)
    assert decided > 50
