"""Readout steps and the fluid mass statistics on the GPU.

A readout step runs the split kernels of a plain step (bulk kernel, listed cells) and also writes the macro
fields (AERO_STEP_MACROS) or the macro fields and the population snapshot (AERO_STEP_READOUT). Its fields must
equal those of the one-cell-per-thread kernel bit for bit. computeFluidMassStats (lbmCpuReference.ts:461-495) is
reduced on the GPU and compared with the oracle's restatement on the same fields: counts, maxima and the NaN cut-off
exactly, the FP64 sums to 1e-12 relative (the GPU sums in a tree, the reference in index order).
"""
import numpy as np
import pytest

import aerosim_amd as A
from aerosim_amd import lbm as L
from tests.helpers import rocket_case, sphere_case

pytestmark = pytest.mark.gpu


def _ctx(storage=A.AERO_STORE_FP32, case="sphere", cpt=None):
    if case == "sphere":
        tris, H = sphere_case(64, 32, 32)
        dims = (64, 32, 32)
    else:
        tris, H = rocket_case(128, 64, 64)
        dims = (128, 64, 64)
    ctx = A.Context(*dims, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(dims[0]), float(dims[2]), H))
    ctx.set_params(A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17))
    ctx.reset(11, 1e-4)
    if cpt is not None:
        ctx.set_cells_per_thread(cpt)
    return ctx


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
@pytest.mark.parametrize("case", ["sphere", "rocket"])
def test_split_readout_fields_equal_per_cell_kernel(storage, case):
    ref = _ctx(storage, case, cpt=0)  # every step through the one-cell-per-thread kernel
    full = _ctx(storage, case)
    mac = _ctx(storage, case)
    for c in (ref, full):
        c.step(9, readout=True)
    mac.step(9, macros=True)
    for u, v, w in zip(ref.read_macro(), full.read_macro(), mac.read_macro()):
        assert np.array_equal(u, v) and np.array_equal(u, w)
    assert np.array_equal(ref.read_populations(), full.read_populations())
    with pytest.raises(A.AeroError) as e:
        mac.read_populations()  # a macros-only step leaves no snapshot
    assert e.value.code == L.AERO_E_STATE
    # later steps are unaffected by the readout
    for c in (ref, full, mac):
        c.step(6, readout=True)
    assert np.array_equal(ref.read_populations(), mac.read_populations())


def test_solid_macros_after_a_new_geometry():
    """Solid cells' macro fields (0, 0, 0, 1) follow a geometry change; the split step never visits them."""
    ctx = _ctx()
    ctx.step(2, macros=True)
    tris, H = sphere_case(64, 32, 32, diameter=12.0, x_centre=30.0)
    ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    ctx.reset(11, 1e-4)
    ctx.step(2, macros=True)
    flags = ctx.read_voxels().cell_flags
    ux, uy, uz, rho = ctx.read_macro()
    s = flags == A.CELL_SOLID
    assert s.sum() > 0
    assert np.all(ux[s] == 0) and np.all(uy[s] == 0) and np.all(uz[s] == 0) and np.all(rho[s] == 1)
    ref = _ctx(cpt=0)
    ref.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    ref.reset(11, 1e-4)
    ref.step(2, macros=True)
    for u, v in zip(ctx.read_macro(), ref.read_macro()):
        assert np.array_equal(u, v)


def _check_stats(ctx):
    from oracle import oracle as O

    g = ctx.mass_stats()
    flags = ctx.read_voxels().cell_flags
    ux, uy, uz, rho = ctx.read_macro()
    o = O.mass_stats_fields(flags, rho, ux, uy, uz)
    print("gpu", g, "\noracle", o)
    assert g["fluid_count"] == o["fluid_count"] and g["has_nan"] == o["has_nan"]
    assert g["max_speed"] == o["max_speed"]
    assert abs(g["total_mass"] - o["total_mass"]) <= 1e-12 * max(1.0, abs(o["total_mass"]))
    assert abs(g["mean_rho"] - o["mean_rho"]) <= 1e-12
    return g


@pytest.mark.parametrize("case", ["sphere", "rocket"])
def test_mass_stats_match_the_oracle_on_the_same_fields(case):
    ctx = _ctx(case=case)
    ctx.step(40, macros=True)
    g = _check_stats(ctx)
    assert not g["has_nan"] and abs(g["mean_rho"] - 1.0) < 1e-3


def test_mass_stats_stop_at_the_first_nan():
    """A NaN planted in one population of one cell reaches the cell it streams to; the statistics cover the
    non-solid cells before the first NaN cell in index order, as the reference's loop does."""
    ctx = _ctx()
    ctx.step(4, readout=True)
    f = ctx.read_populations().reshape(19, -1)
    idx = 40 + 64 * (16 + 32 * 16)  # a fluid cell downstream of the sphere, mid-tunnel
    assert ctx.read_voxels().cell_flags[idx] == A.CELL_FLUID
    f[5, idx] = np.nan
    ctx.write_populations(f.reshape(-1))
    ctx.step(1, macros=True)
    g = _check_stats(ctx)
    # the planted +z population streams to the cell above (z + 1): the first NaN cell in index order
    flags = ctx.read_voxels().cell_flags
    first = idx + 64 * 32
    assert g["has_nan"] and g["fluid_count"] == int(np.count_nonzero(flags[:first] != A.CELL_SOLID))
