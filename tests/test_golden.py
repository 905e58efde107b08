"""Regression fixtures of tests/golden (written by tests/golden/make_fixtures.py from the oracle)."""
import os

import numpy as np
import pytest

import aerosim_amd as A

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "sphere16.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX, allow_pickle=False))


def test_mesh_and_voxels_reproduce(fx, oracle_lib):
    from tests.golden.make_fixtures import build_case

    tris, vox = build_case()
    assert np.array_equal(tris, fx["tris"])
    assert np.array_equal(vox.cell_flags, fx["flags"])
    assert np.array_equal(vox.boundary_q_values.view(np.uint32), fx["q"].view(np.uint32))


@pytest.mark.parametrize("mode", ["fp64", "ref32"])
def test_oracle_reproduces_fixture(fx, oracle_lib, mode):
    from tests.golden.make_fixtures import CASE, build_case

    O = oracle_lib
    c = CASE
    _, vox = build_case()
    s = O.Solver(vox, mode, u_inlet=c["u"], nu=c["nu"], cs=c["cs"])
    s.reset(c["seed"], c["amp"])
    s.step(1)
    f1 = s.populations()
    s.step(9)
    f10 = s.populations()
    # transcendental library calls (sin/pow in the inlet rule) may differ by an ulp between hosts
    assert np.max(np.abs(f1 - fx[f"f1_{mode}"])) < 1e-12
    assert np.max(np.abs(f10 - fx[f"f10_{mode}"])) < 1e-12


@pytest.mark.gpu
def test_gpu_matches_fixture(fx):
    from tests.golden.make_fixtures import CASE

    c = CASE
    ctx = A.Context(c["nx"], c["ny"], c["nz"])
    ctx.set_mesh(fx["tris"], A.Tunnel(float(c["nx"]), float(c["nz"]), float(c["ny"])))
    gv = ctx.read_voxels(float(c["ny"]))
    assert np.array_equal(gv.cell_flags, fx["flags"])
    ctx.set_params(A.SolverParams(u_inlet=c["u"], nu=c["nu"], smagorinsky_cs=c["cs"]))
    ctx.reset(c["seed"], c["amp"])
    ctx.step(10, readout=True)
    m = fx["flags"] != A.CELL_SOLID
    fg = ctx.read_populations().reshape(19, -1)[:, m].astype(np.float64)
    fo = fx["f10_fp64"].reshape(19, -1)[:, m]
    err = np.max(np.abs(fg - fo) / np.abs(fo))
    print(f"gpu vs fixture (fp64 oracle, 10 steps): {err:.2e}")
    assert err <= 1e-5
    f = ctx.read_forces()
    fo_forces = fx["forces10_fp64"]
    assert abs(f["fx"] - fo_forces[0]) <= 1e-4 * abs(fo_forces[0]) + 1e-7
