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


def test_closure_forward_in_the_solver_oracle_equals_the_numpy_network(oracle_lib):
    """The C++ oracle's network (used inside its collision) equals mlp_oracle.forward with the same BF16 operand
    rounding, for seeded parameters and features spread over the features' ranges."""
    rng = np.random.default_rng(3)
    p = M.init_params(seed=4, scale=0.4).astype(np.float32)
    X = np.zeros((500, 8))
    X[:, :6] = rng.uniform(-1.5, 1.5, (500, 6))
    y_cpp = oracle_lib.closure_forward(p, X)
    y_np, _ = M.forward(p.astype(np.float64), X, bf=True)
    assert np.max(np.abs(y_cpp - y_np)) < 1e-12


def test_closure_features_closed_forms(oracle_lib):
    """Isotropic stress: tr/|P| = sqrt 3, no deviator; pure shear: trace 0, all deviatoric, shape 0; uniaxial
    deviator diag(2, -1, -1): shape 1 (its negative: -1); magnitude and viscosity features ln(.)/10 + 1; the
    Smagorinsky feature ln(1 + nu_S / nu) / 4 with tau_S = (tau0 + sqrt(tau0^2 + smag |P|)) / 2."""
    F = oracle_lib.closure_features
    nu, tau0, smag = 0.004, 0.512, 18 * np.sqrt(2) * 0.17 ** 2
    x = F([2e-3, 2e-3, 2e-3, 0, 0, 0], nu, tau0, smag)
    n = np.sqrt(3) * 2e-3
    assert abs(x[0] - (np.log(n + 1e-12) / 10 + 1)) < 1e-12 and abs(x[1] - np.sqrt(3)) < 1e-9
    assert abs(x[2]) < 1e-12 and abs(x[3]) < 1e-9
    assert abs(x[4] - (np.log(nu) / 10 + 1)) < 1e-12
    tau_s = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + smag * n))
    assert abs(x[5] - np.log1p((tau_s - tau0) / 3 / nu) / 4) < 1e-12
    x = F([0, 0, 0, 1e-3, 0, 0], nu, tau0, smag)
    assert abs(x[1]) < 1e-12 and abs(x[2] - 1.0) < 1e-12 and abs(x[3]) < 1e-9
    assert abs(F([2e-3, -1e-3, -1e-3, 0, 0, 0], nu, tau0, smag)[3] - 1.0) < 1e-9
    assert abs(F([-2e-3, 1e-3, 1e-3, 0, 0, 0], nu, tau0, smag)[3] + 1.0) < 1e-9
    # invariants: a rotation of the stress leaves every feature unchanged
    P = np.array([[3e-4, 1e-4, -2e-4], [1e-4, -1e-4, 5e-5], [-2e-4, 5e-5, 4e-4]])
    c, s = np.cos(0.7), np.sin(0.7)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]) @ np.array([[1, 0, 0], [0, c, s], [0, -s, c]])
    Q = R @ P @ R.T
    v = lambda A: [A[0, 0], A[1, 1], A[2, 2], A[0, 1], A[0, 2], A[1, 2]]
    assert np.allclose(F(v(P), nu, tau0, smag), F(v(Q), nu, tau0, smag), atol=1e-9)


def test_zero_network_closure_is_smagorinsky_in_the_oracle(oracle_lib):
    """With all parameters zero the factor is softplus(ln(e - 1)) = 1, so the closure's step equals the
    Smagorinsky step (to rounding); a network whose output is +2 everywhere (b3 = 2) raises tau and so changes it."""
    O = oracle_lib
    import aerosim_amd as A

    tris = A.mesh_icosphere(2, 3.0, (10.0 - 16.0, 0.0, 0.0))
    vox = O.Voxels(tris, 16.0, (32, 16, 16))
    runs = []
    for prm in (None, np.zeros(M.NPARAM, np.float32), np.r_[np.zeros(M.NPARAM - 1), 2.0].astype(np.float32)):
        s = O.Solver(vox, "fp64", u_inlet=0.05, nu=0.004, cs=0.17)
        s.set_closure(prm)
        s.reset(1, 1e-4)
        s.step(6)
        runs.append(s.populations())
    assert np.max(np.abs(runs[1] - runs[0])) < 1e-15
    assert np.max(np.abs(runs[2] - runs[0])) > 1e-9


def test_sgs_samples_of_a_linear_field():
    """u = A x on the fine grid: 2 x 2 x 2 block means are the field at the block centres, their central
    differences give the gradient 2 A per coarse cell, and the sub-filter stress over a block is
    tau_ac = sum_b A_ab A_cb / 4 (each fine coordinate takes two values one apart: variance 1/4), so every sample
    carries the closed-form nu_r = -dev(tau) : S / (2 S : S) with S = A + A^T (per coarse cell)."""
    fx, fy, fz = 16, 12, 10
    A = np.array([[1e-3, 4e-4, -2e-4], [3e-4, -6e-4, 1e-4], [-1e-4, 2e-4, -4e-4]])
    z, y, x = np.meshgrid(np.arange(fz), np.arange(fy), np.arange(fx), indexing="ij")
    pos = np.stack([x, y, z]).astype(np.float64)
    u = np.einsum("ab,b...->a...", A, pos) + np.array([0.05, 0.0, 0.0])[:, None, None, None]
    nu, cs = 0.002, 0.17
    X, T, det = M.sgs_samples(u[0].ravel(), u[1].ravel(), u[2].ravel(), np.zeros(fx * fy * fz, np.uint8),
                              (fx, fy, fz), 1, 0, nu, cs)
    assert len(det) == (fx // 2 - 2) * (fy // 2 - 2) * (fz // 2 - 2)
    S = A + A.T
    tau = A @ A.T / 4
    tau -= np.eye(3) * np.trace(tau) / 3
    nu_r = -np.sum(tau * S) / (2 * np.sum(S * S))
    smag = 18 * np.sqrt(2) * cs * cs
    nut = nu_r + 0.25 * smag / 18 * np.sqrt(np.sum(S * S))
    for d in det:
        assert np.allclose(d["S"], S, rtol=1e-9, atol=1e-15) and abs(d["nut"] - nut) < 1e-12
    tau_t = 3 * (nu + max(0.0, nut)) + 0.5
    P = -(2 / 3) * tau_t * S
    Xe = M.closure_features(np.array([[P[0, 0], P[1, 1], P[2, 2], P[0, 1], P[0, 2], P[1, 2]]]), nu, 3 * nu + 0.5, smag)
    assert np.allclose(X[:, :6], Xe, rtol=1e-9, atol=1e-9)
    nu_s = det[0]["nu_s"]
    assert np.allclose(T, nut / nu_s, rtol=1e-9)
    # a solid cell removes every sample whose own or neighbouring block holds it
    fl = np.zeros((fz, fy, fx), np.uint8)
    fl[4, 6, 8] = 1  # block (i, j, k) = (4, 3, 2)
    _, _, det2 = M.sgs_samples(u[0].ravel(), u[1].ravel(), u[2].ravel(), fl.ravel(), (fx, fy, fz), 1, 0, nu, cs)
    gone = {d["ijk"] for d in det} - {d["ijk"] for d in det2}
    assert gone == {(4, 3, 2), (3, 3, 2), (5, 3, 2), (4, 2, 2), (4, 4, 2), (4, 3, 1), (4, 3, 3)}
