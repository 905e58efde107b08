"""Host-side pieces of the learned-closure workflow (no GPU): the seeded training sizes and the constant-factor
network used by the factor sweep."""
import numpy as np

from aerosim_amd import closure as C
from oracle import mlp_oracle as MO


def test_seeded_sizes_lie_on_cell_boundaries_of_both_grids():
    sizes = C.seeded_sizes(12, seed=100)
    assert len(set(sizes)) == 12 and C.NOMINAL not in sizes
    case = C.BuildingCase()
    x0 = case.x_front * case.nx - 0.5 * case.nx
    for bx, bz, bh in sizes:
        assert all(isinstance(v, int) for v in (bx, bz, bh)) and bz % 2 == 0
        for b, nominal in zip((bx, bz, bh), C.NOMINAL):
            assert abs(b / nominal - 1.0) <= 0.1 + 1.0 / nominal  # +-10 % before rounding to whole cells
        for f in (1, 2):  # every face on a cell boundary of the coarse and of the fine grid
            faces = [f * x0, f * (x0 + bx), f * 0.5 * bz, f * (-0.5 * case.height + bh)]
            assert all(float(v).is_integer() for v in faces)
    assert C.seeded_sizes(4, seed=100) == C.seeded_sizes(4, seed=100) != C.seeded_sizes(4, seed=101)


def test_constant_factor_parameters():
    """factor_params(g) through the oracle's forward pass gives, for any input, the output y whose factor
    softplus(y + ln(e - 1)) is g."""
    X = np.random.default_rng(0).standard_normal((64, 8)).astype(np.float32)
    for g in (0.5, 0.9, 1.0, 2.0):
        y, _ = MO.forward(C.factor_params(g).astype(np.float64), X, bf=False)
        assert np.allclose(np.log1p(np.exp(y + np.log(np.e - 1.0))), g, rtol=1e-6, atol=0)


def test_strain_gate_parameters():
    """gate_params(g, c) through the oracle's forward pass: factor g well below the threshold c of x5, 1 well above,
    and output y linear in x5 across the ramp (BF16 operands: to 1e-2)."""
    g, c = 0.62, 0.9
    X = np.zeros((7, 8), np.float32)
    X[:, 5] = [0.0, 0.5, 0.79, 0.85, 0.9, 1.01, 1.5]
    X[:, :5] = np.random.default_rng(1).standard_normal((7, 5))  # the other features do not matter
    p = C.gate_params(g, c).astype(np.float64)
    y, _ = MO.forward(p, X, bf=True)
    fac = np.log1p(np.exp(y + np.log(np.e - 1.0)))
    assert np.allclose(fac[:3], g, atol=1e-2) and np.allclose(fac[5:], 1.0, atol=1e-2)
    y_lo = np.log(np.expm1(g)) - np.log(np.e - 1.0)
    assert np.allclose(y[3:5], y_lo * (1 - (X[3:5, 5] - (c - 0.1)) / 0.2), atol=1e-2)
    assert C.gate_params(g, c).size == 1313 == MO.NPARAM
