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
    X[::3, 7] = np.random.default_rng(5).uniform(0.2, 4.0, size=X[::3].shape[0])  # loss weights on a third
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


def test_zero_network_is_the_smagorinsky_closure():
    """The zero network's output y = 0 maps to the factor 1 on the Smagorinsky viscosity, and the factor is
    positive and increasing (softplus shifted by ln(e - 1))."""
    assert abs(M.factor(0.0) - 1.0) < 1e-15
    ys = np.linspace(-30, 30, 61)
    g = M.factor(ys)
    assert np.all(g > 0) and np.all(np.diff(g) > 0)
    assert abs(M.factor(25.0) - (25.0 + M.C0)) < 1e-9       # linear for large y
    p = np.zeros(M.NPARAM)
    X, _ = _data(50)
    assert np.allclose(M.predict_factor(p, X), 1.0)


def test_weight_column_never_reaches_the_network():
    """Column 7 carries the loss weight: the forward pass ignores it, and the loss scales with it."""
    p = M.init_params(4)
    X, T = _data(64)
    y0, _ = M.forward(p, X, bf=False)
    X2 = X.copy()
    X2[:, 7] = 3.0
    y1, _ = M.forward(p, X2, bf=False)
    assert np.array_equal(y0, y1)
    l0, _ = M.loss_and_grad(p, X, T, bf=False)
    l1, _ = M.loss_and_grad(p, X2, T, bf=False)
    assert abs(l1 - 3.0 * l0) <= 1e-12 * abs(l1)
