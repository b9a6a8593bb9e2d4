"""Pin the numpy MLP oracle (oracle/mlp_oracle.py): its analytic gradient
equals central finite differences of its loss."""
import numpy as np

from oracle import mlp_oracle as M


def test_parameter_count_matches_layout():
    assert M.P == 6 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 8833


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.2, M.P)
    X = rng.normal(size=(17, 6))
    T = rng.normal(size=17)
    _, g = M.loss_and_grad(p, X, T)
    idx = rng.choice(M.P, 60, replace=False)
    idx = np.concatenate([idx, [M.P - 1, 0, 6 * 64]])  # b4, W1[0,0], b1[0]
    for i in idx:
        e = np.zeros(M.P)
        e[i] = 1e-6
        fd = (M.loss_and_grad(p + e, X, T)[0] - M.loss_and_grad(p - e, X, T)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i
