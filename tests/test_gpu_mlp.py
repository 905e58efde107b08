"""GPU checks of the learned closure: matrix-core forward/backward against the CPU oracle, training, and the
closure fused into the collision."""
import numpy as np
import pytest

import aerosim_amd as A
from oracle import mlp_oracle as MO
from tests.helpers import relerr, sphere_case

pytestmark = pytest.mark.gpu


def _data(n, seed=1):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 8), dtype=np.float32)
    X[:, :6] = (rng.standard_normal((n, 6)) * 0.7).astype(np.float32)
    T = (np.abs(rng.standard_normal(n)) * 0.5).astype(np.float32)
    return X, T


def test_forward_matches_oracle():
    p = MO.init_params(5).astype(np.float32)
    m = A.Closure(params=p)
    X, _ = _data(1000)
    y = m.predict(X)
    yo, _ = MO.forward(p.astype(np.float64), X, bf=True)
    err = np.max(np.abs(y - yo)) / (np.max(np.abs(yo)) + 1e-12)
    print(f"forward max err vs oracle (bf16 operands emulated): {err:.2e}")
    assert err < 1e-4


def test_gradient_matches_oracle():
    p = MO.init_params(7).astype(np.float32)
    m = A.Closure(params=p)
    X, T = _data(777)  # not a multiple of 64: masked lanes must not contribute
    X[::2, 7] = np.random.default_rng(3).uniform(0.1, 5.0, size=X[::2].shape[0])  # loss weights on half
    loss, g = m.grad(X, T)
    lo, go = MO.loss_and_grad(p.astype(np.float64), X, T, bf=True)
    print(f"loss gpu {loss:.8f} oracle {lo:.8f}; grad rel err {np.max(np.abs(g - go)) / np.max(np.abs(go)):.2e}")
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert np.max(np.abs(g - go)) <= 2e-3 * np.max(np.abs(go))
    # against the exact FP64 gradient the BF16 operands cost more, but stay small
    _, ge = MO.loss_and_grad(p.astype(np.float64), X, T, bf=False)
    assert np.max(np.abs(g - ge)) <= 5e-2 * np.max(np.abs(ge))


def test_training_reduces_loss():
    m = A.Closure(params=MO.init_params(1).astype(np.float32))
    rng = np.random.default_rng(2)
    X = np.zeros((32768, 8), dtype=np.float32)
    X[:, :6] = rng.standard_normal((32768, 6)).astype(np.float32)
    T = np.log1p(np.abs(X[:, 0] * 2 + X[:, 5])).astype(np.float32)  # a smooth positive target
    m.add_samples(X, T)
    losses = m.train(400, batch=4096, lr=3e-3)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.2 * losses[0]


def test_zero_network_equals_smagorinsky():
    """The zero network (factor 1 on the Smagorinsky viscosity) run through the fused matrix-core closure gives
    the populations of the plain Smagorinsky step."""
    nu = 0.004
    m = A.Closure(params=np.zeros(A.lbm.MLP_NPARAM, dtype=np.float32))
    tris, H = sphere_case(64, 32, 32)
    runs = []
    for mode in ("mlp", "smag"):
        ctx = A.Context(64, 32, 32)
        ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
        if mode == "mlp":
            ctx.set_closure(m)
            ctx.set_params(A.SolverParams(u_inlet=0.05, nu=nu, smagorinsky_cs=0.17, closure=A.lbm.AERO_CLOSURE_MLP))
        else:
            ctx.set_params(A.SolverParams(u_inlet=0.05, nu=nu, smagorinsky_cs=0.17))
        ctx.reset(3, 1e-4)
        ctx.step(20, readout=True)
        runs.append((ctx, ctx.read_populations()))
    flags = runs[0][0].read_voxels().cell_flags
    mm = flags != A.CELL_SOLID
    err = relerr(runs[0][1].reshape(19, -1)[:, mm], runs[1][1].reshape(19, -1)[:, mm])
    print(f"zero-network closure run vs Smagorinsky run: {err:.2e}")
    assert err < 1e-5


def test_closure_factor_raises_dissipation():
    """A constant factor above 1 (b3 > 0) raises the eddy viscosity: seeded velocity noise in quiescent fluid
    decays faster than with the plain Smagorinsky closure, and a factor below 1 makes it decay slower."""
    energies = {}
    for name, b3 in (("smagorinsky", None), ("factor 3.6", 3.0), ("factor 0.1", -2.8)):
        tris, H = sphere_case(64, 32, 32)
        ctx = A.Context(64, 32, 32)
        ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
        if b3 is not None:
            p = np.zeros(A.lbm.MLP_NPARAM, dtype=np.float32)
            p[-1] = b3
            ctx.set_closure(A.Closure(params=p))
        ctx.set_params(A.SolverParams(u_inlet=0.0, nu=0.0005, smagorinsky_cs=0.17,
                                      closure=A.lbm.AERO_CLOSURE_MLP if b3 is not None else A.lbm.AERO_CLOSURE_SMAGORINSKY))
        ctx.reset(11, 2e-2)
        ctx.step(40, readout=True)
        ux, uy, uz, _ = ctx.read_macro()
        fl = ctx.read_voxels().cell_flags != A.CELL_SOLID
        energies[name] = float(np.mean((ux * ux + uy * uy + uz * uz)[fl]))
    print("kinetic energy after 40 steps:", energies)
    assert energies["factor 3.6"] < energies["smagorinsky"] < energies["factor 0.1"]


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_split_closure_step_matches_per_cell_step(storage):
    """Plain steps with the learned closure run split (bulk cells, then the listed cells) and give the
    populations of the per-cell closure kernel (cells per thread 0), bit for bit: a sample's network output
    does not depend on which cells share its wavefront. A seeded random network (so the factor varies from
    cell to cell), the sphere plus one-cell plates in the inlet and outlet planes."""
    rng = np.random.default_rng(5)
    m = A.Closure(params=(rng.standard_normal(A.lbm.MLP_NPARAM) * 0.3).astype(np.float32))
    tris, H = sphere_case(64, 32, 32)
    plates = [A.mesh_box((-33.0, -4.0, -3.0), (-31.2, 4.0, 5.0)), A.mesh_box((31.2, -6.0, -5.0), (33.0, 2.0, 3.0))]
    pops = []
    for split in (True, False):
        ctx = A.Context(64, 32, 32, storage=storage)
        ctx.set_mesh(np.concatenate([tris] + plates), A.Tunnel(64.0, 32.0, H))
        ctx.set_closure(m)
        ctx.set_params(A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17, closure=A.lbm.AERO_CLOSURE_MLP))
        ctx.reset(3, 1e-4)
        if split:
            ctx.step(13, readout=True)  # twelve split steps, then one readout step
            info = ctx.split_info()
            print("split step with the closure:", info)
            assert info["cells_per_thread"] == 1
        else:
            ctx.set_cells_per_thread(0)  # every step through the per-cell closure kernel
            ctx.step(13, readout=True)
        pops.append(ctx.read_populations())
        ctx.close()
    print(f"max |split - per-cell| {float(np.abs(pops[0] - pops[1]).max()):.3e}")
    assert np.array_equal(pops[0], pops[1])
