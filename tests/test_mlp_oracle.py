"""Pins of the learned-closure oracle (oracle/mlp_oracle.py): its analytic gradient against central finite
differences of the loss, and the feature definition against closed forms."""
import numpy as np

from oracle import mlp_oracle as M


def _data(n=200, seed=1):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 8))
    X[:, :6] = rng.standard_normal((n, 6)) * 0.7
    T = np.abs(rng.standard_normal(n)) * 0.5
    return X, T


def test_gradient_matches_finite_differences():
    p = M.init_params(3)
    X, T = _data()
    _, g = M.loss_and_grad(p, X, T, bf=False)
    rng = np.random.default_rng(0)
    idx = rng.choice(M.NPARAM, 80, replace=False)
    h = 1e-6
    for j in idx:
        e = np.zeros(M.NPARAM)
        e[j] = h
        lp, _ = M.loss_and_grad(p + e, X, T, bf=False)
        lm, _ = M.loss_and_grad(p - e, X, T, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (j, fd, g[j])


def test_bf16_rounding_is_round_to_nearest_even():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2 ** -8) == 1.0          # halfway, rounds to even
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # halfway, rounds up to even
    assert M.bf16(1.0 + 2 ** -7) == 1.0 + 2 ** -7


def test_parameter_count():
    assert M.NPARAM == 1313
