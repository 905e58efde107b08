"""GPU parity of the HIP step path against the CPU oracle (same seeded inputs, same voxels).

Tolerances (stated by the workload definition): populations within 1e-5 relative in FP32, Cd within
0.5 %, voxel flags bit-exact. Every check prints its error against the FP64 oracle.
"""
import numpy as np
import pytest

import aerosim_amd as A
from tests.helpers import relerr, rocket_case, sphere_case

pytestmark = pytest.mark.gpu

POP_TOL = 1e-5


def _pair(vox, mode="fp64", **kw):
    from oracle import oracle as O

    p = dict(u_inlet=0.05, nu=0.004, cs=0.17)
    p.update(kw)
    orc = O.Solver(vox, mode, u_inlet=p["u_inlet"], nu=p["nu"], cs=p["cs"], ground_noslip=p.get("noslip", False),
                   profile=p.get("profile", 0), z0=p.get("z0", 0.5), zref=p.get("zref", 64.0))
    ctx = A.Context(vox.nx, vox.ny, vox.nz, storage=p.get("storage", A.AERO_STORE_FP32))
    ctx.set_voxels(vox)
    ctx.set_params(A.SolverParams(u_inlet=p["u_inlet"], nu=p["nu"], smagorinsky_cs=p["cs"],
                                  ground_noslip=p.get("noslip", False), inflow_profile=p.get("profile", 0),
                                  z0=p.get("z0", 0.5), z_ref=p.get("zref", 64.0)))
    return orc, ctx


def _nonsolid(vox):
    return np.asarray(vox.cell_flags) != A.CELL_SOLID


def _compare_pops(orc, ctx, vox, tol, label):
    fo = orc.populations().reshape(19, -1)
    fg = ctx.read_populations().reshape(19, -1).astype(np.float64)
    m = _nonsolid(vox)
    err = relerr(fg[:, m], fo[:, m])
    print(f"[{label}] max rel err populations vs oracle = {err:.3e}")
    assert np.all(np.isfinite(fg[:, m]))
    assert err <= tol, f"{label}: {err}"
    return err


@pytest.fixture(scope="module")
def c1_vox():
    from oracle import oracle as O

    tris, H = sphere_case(64, 32, 32)
    return O.Voxels(tris, H, (64, 32, 32))


def test_init_matches_oracle(c1_vox):
    orc, ctx = _pair(c1_vox)
    orc.reset(1234, 1e-4)
    ctx.reset(1234, 1e-4)
    ctx.step(0)
    # state after reset: step 0 readout is not defined; step once and compare instead in the next test,
    # here compare the forces of the initial state (uses only the stored populations)
    fo = orc.forces()
    fg = ctx.read_forces()
    for k in ("fx", "fy", "fz", "mx", "my", "mz"):
        assert abs(fg[k] - fo[k]) <= 1e-6 * max(1.0, abs(fo[k])), (k, fg[k], fo[k])
    assert fg["cut_face_links"] == fo["cut_face_links"]


@pytest.mark.parametrize("nsteps", [1, 2, 10])
def test_populations_sphere(c1_vox, nsteps):
    orc, ctx = _pair(c1_vox)
    orc.reset(1234, 1e-4)
    ctx.reset(1234, 1e-4)
    orc.step(nsteps)
    ctx.step(nsteps, readout=True)
    _compare_pops(orc, ctx, c1_vox, POP_TOL, f"sphere {nsteps} steps")
    ux, uy, uz, rho = ctx.read_macro()
    oux, ouy, ouz, orho = orc.macro()
    m = _nonsolid(c1_vox)
    assert relerr(rho[m], orho[m]) <= POP_TOL
    assert np.max(np.abs(ux[m] - oux[m])) <= 1e-6
    assert np.max(np.abs(uy[m] - ouy[m])) <= 1e-6
    fo, fg = orc.forces(), ctx.read_forces()
    print(f"[sphere {nsteps}] fx gpu {fg['fx']:.8f} oracle {fo['fx']:.8f}")
    assert abs(fg["fx"] - fo["fx"]) <= 1e-4 * abs(fo["fx"]) + 1e-7


def test_ref32_mode_close_to_fp64(c1_vox):
    """The reference's own FP32 storage error (ref32 vs fp64) and the GPU error are the same order."""
    o64, ctx = _pair(c1_vox)
    o32, _ = _pair(c1_vox, "ref32")
    for o in (o64, o32):
        o.reset(7, 1e-4)
        o.step(10)
    ctx.reset(7, 1e-4)
    ctx.step(10, readout=True)
    m = _nonsolid(c1_vox)
    f64 = o64.populations().reshape(19, -1)[:, m]
    e32 = relerr(o32.populations().reshape(19, -1)[:, m], f64)
    eg = relerr(ctx.read_populations().reshape(19, -1)[:, m], f64)
    print(f"ref32 vs fp64 {e32:.3e}; gpu vs fp64 {eg:.3e}")
    assert eg <= POP_TOL and e32 <= POP_TOL


def test_single_step_from_injected_state(c1_vox):
    """One step from the oracle's state after 50 steps (isolates the per-step error)."""
    orc, ctx = _pair(c1_vox)
    orc.reset(99, 1e-4)
    orc.step(50)
    ctx.reset(99, 1e-4)
    ctx.write_populations(orc.populations().astype(np.float32))
    ctx.set_step_index(orc.step_index)
    # the oracle continues from its double state; give it the same float-rounded state
    orc.set_populations(orc.populations().astype(np.float32).astype(np.float64))
    orc.step(1)
    ctx.step(1, readout=True)
    _compare_pops(orc, ctx, c1_vox, POP_TOL, "step 51 from injected state")


def test_zero_inflow_force(c1_vox):
    """phase4.test.ts:11-50: zero inflow gives |F| < 1e-6."""
    _, ctx = _pair(c1_vox, u_inlet=0.0, nu=0.005, cs=0.12)
    ctx.reset(0, 0.0)
    ctx.step(5)
    f = ctx.read_forces()
    assert abs(f["fx"]) < 1e-6 and abs(f["fy"]) < 1e-6 and abs(f["fz"]) < 1e-6


def test_mass_anchor_fires_and_matches(c1_vox):
    """Scale all populations by 1.001 so the global mass anchor fires on the first step."""
    orc, ctx = _pair(c1_vox)
    orc.reset(5, 1e-4)
    f = (orc.populations() * 1.001).astype(np.float32)
    orc.set_populations(f.astype(np.float64))
    ctx.reset(5, 1e-4)
    ctx.write_populations(f)
    ctx.set_step_index(orc.step_index)
    orc.step(3)
    ctx.step(3, readout=True)
    assert orc.anchor_count() >= 1
    assert ctx.anchor_count() == orc.anchor_count()
    _compare_pops(orc, ctx, c1_vox, POP_TOL, "anchor")


@pytest.mark.parametrize("mode", ["log law, no-slip", "power law, no-slip", "uniform, free slip"])
def test_ground_building(mode):
    """Block on the floor (boundary cells on the floor face, which the wall rules skip) under the three floor and
    inflow modes: the log law of the boundary-layer workload, the reference's own ground mode (power-law inflow,
    no-slip floor, lbmCpuReference.ts:122-125, 335-338, 378-390) and the free-slip floor."""
    from oracle import oracle as O

    nx, ny, nz = 48, 24, 32
    tris = A.mesh_box((-6.0, -ny / 2 - 1.0, -3.0), (2.0, -ny / 2 + 9.0, 3.0))
    vox = O.Voxels(tris, float(ny), (nx, ny, nz))
    kw = {"log law, no-slip": dict(noslip=True, profile=2, z0=0.5, zref=12.0),
          "power law, no-slip": dict(noslip=True, profile=0), "uniform, free slip": dict(noslip=False, profile=0)}[mode]
    orc, ctx = _pair(vox, u_inlet=0.06, nu=0.002, cs=0.17, **kw)
    orc.reset(3, 1e-4)
    ctx.reset(3, 1e-4)
    orc.step(10)
    ctx.step(10, readout=True)
    _compare_pops(orc, ctx, vox, POP_TOL, f"building/{mode}")
    fo, fg = orc.forces(), ctx.read_forces()
    assert abs(fg["fx"] - fo["fx"]) <= 1e-4 * abs(fo["fx"]) + 1e-7


def test_quasi2d_cylinder_no_sides():
    """nz = 3: no side walls, clamped pulls in z on every cell (phase3.test.ts:45-124 geometry class)."""
    from oracle import oracle as O

    nx, ny, nz = 48, 20, 3
    tris = A.mesh_naca(0.99 * 0.0 + 0.3, 6.0, (-12.0, 0.0, 0.0), 0.0, -4.0, 4.0, 32)
    vox = O.Voxels(tris, float(ny), (nx, ny, nz))
    orc, ctx = _pair(vox, u_inlet=0.08, nu=0.003, cs=0.12)
    orc.reset(11, 1e-4)
    ctx.reset(11, 1e-4)
    orc.step(10)
    ctx.step(10, readout=True)
    _compare_pops(orc, ctx, vox, POP_TOL, "quasi-2d")


def test_rocket_downscale_parity():
    from oracle import oracle as O

    tris, H = rocket_case(128, 64, 64)
    vox = O.Voxels(tris, H, (128, 64, 64))
    orc, ctx = _pair(vox, u_inlet=0.08, nu=0.08 * 6 / 2e4 * 4, cs=0.17)
    orc.reset(2024, 1e-4)
    ctx.reset(2024, 1e-4)
    orc.step(5)
    ctx.step(5, readout=True)
    _compare_pops(orc, ctx, vox, POP_TOL, "rocket 128x64x64")


def test_ramp_protocol(c1_vox):
    """sin^2 ramp of testRunner.ts:89-96 computed on the device vs the host-driven oracle."""
    orc, ctx = _pair(c1_vox, u_inlet=0.05)
    ctx.set_params(A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17, ramp_steps=80, u_init=0.001))
    orc.set_u_inlet(0.001)
    orc.reset(0, 0.0)
    ctx.reset(0, 0.0)
    for s in range(1, 13):
        frac = s / 80
        orc.set_u_inlet(0.05 * np.sin(0.5 * np.pi * frac) * np.sin(0.5 * np.pi * frac))
        orc.step(1)
    ctx.step(12, readout=True)
    _compare_pops(orc, ctx, c1_vox, POP_TOL, "ramp")


def test_cd_1000_steps_c1(c1_vox):
    """Config 1 in full: Cd (mean over the last 40 samples of the second half) within 0.5 % of FP64."""
    from oracle import oracle as O

    u, D, Re = 0.05, 8.0, 100.0
    nu = u * D / Re
    orc, ctx = _pair(c1_vox, u_inlet=u, nu=nu, cs=0.17)
    orc.reset(1234, 1e-4)
    ctx.reset(1234, 1e-4)
    tris, _ = sphere_case()
    st = A.mesh_stats(tris)
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=32.0,
                tunnel_height=32.0, dx=1.0, u_ref=u)
    cd_o, cd_g = [], []
    steps = 1000
    for s in range(1, steps + 1):
        orc.step(1)
        ctx.step(1)
        if s > steps // 2 and s > steps - 40:
            fo = orc.forces()
            cd_o.append(O.coefficients(fo, refs["frontal_area"], refs["char_length"], 32.0, 32.0, 1.0, u)["cd"])
            cd_g.append(ctx.read_forces(refs)["cd"])
    mo, mg = float(np.mean(cd_o)), float(np.mean(cd_g))
    print(f"C1 Cd: gpu {mg:.6f} fp64 oracle {mo:.6f} |dCd| {abs(mg - mo) / abs(mo) * 100:.4f} %")
    assert abs(mg - mo) <= 0.005 * abs(mo)


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_cl_cd_lifting_airfoil(storage):
    """A lifting body: extruded NACA 0012 (chord 24 cells) at 8 degrees in a 96 x 48 x 16 tunnel with side walls,
    u = 0.06, Re 500 on the chord. Mean Cd, Cl and Cm over steps 601-1000 within 0.5 % of the FP64 oracle (Cl and Cm
    being what the symmetric bodies leave near zero); in FP32 and in the shifted-FP16 storage. The centre of pressure
    is printed only: the reference takes the moment about the domain centre (forceIntegrator.ts:214-216), which lies
    0.3 nx behind the airfoil, so -Mz/Fy clamps to half the body length in both runs."""
    from oracle import oracle as O

    nx, ny, nz = 96, 48, 16
    chord, alpha, u, re = 24.0, 8.0, 0.06, 500.0
    tris = A.mesh_naca(0.12, chord, (-0.3 * nx, 0.0, 0.0), alpha, -nz, nz, 48)
    vox = O.Voxels(tris, float(ny), (nx, ny, nz))
    orc, ctx = _pair(vox, u_inlet=u, nu=u * chord / re, cs=0.17, storage=storage)
    orc.reset(7, 1e-4)
    ctx.reset(7, 1e-4)
    st = A.mesh_stats(tris)
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=float(nz),
                tunnel_height=float(ny), dx=1.0, u_ref=u)
    orc.step(600)
    ctx.step(600)
    go, gg = [], []
    for _ in range(400):
        orc.step(1)
        ctx.step(1)
        c = O.coefficients(orc.forces(), refs["frontal_area"], refs["char_length"], float(nz), float(ny), 1.0, u)
        go.append((c["cd"], c["cl"], c["cm"], c["xcp"]))
        f = ctx.read_forces(refs)
        gg.append((f["cd"], f["cl"], f["cm"], f["xcp"]))
    (cdo, clo, cmo, xo), (cdg, clg, cmg, xg) = np.mean(go, axis=0), np.mean(gg, axis=0)
    print(f"NACA 0012 at 8 deg, {'fp16' if storage else 'fp32'} storage: Cd gpu {cdg:.6f} fp64 {cdo:.6f} "
          f"({abs(cdg / cdo - 1) * 100:.4f} %); Cl gpu {clg:.6f} fp64 {clo:.6f} ({abs(clg / clo - 1) * 100:.4f} %); "
          f"Cm gpu {cmg:.6f} fp64 {cmo:.6f} ({abs(cmg / cmo - 1) * 100:.4f} %); centre of pressure gpu {xg:.4f} fp64 "
          f"{xo:.4f} cells")
    assert clo > 0.1  # a lifting case
    assert abs(cdg - cdo) <= 0.005 * abs(cdo)
    assert abs(clg - clo) <= 0.005 * abs(clo)
    assert abs(cmg - cmo) <= 0.005 * abs(cmo)


def test_gpu_voxelizer_bit_exact_sphere():
    from oracle import oracle as O

    tris, H = sphere_case(64, 32, 32)
    ov = O.Voxels(tris, H, (64, 32, 32))
    gv = A.voxelize_mesh_to_lattice(tris, None, A.Tunnel(64.0, 32.0, H), (64, 32, 32))
    assert gv.solid_cell_count == ov.solid_cell_count
    assert gv.boundary_cell_count == ov.boundary_cell_count
    assert np.array_equal(gv.cell_flags, ov.cell_flags)
    assert np.array_equal(gv.boundary_cell_indices, ov.boundary_cell_indices)
    assert np.array_equal(gv.boundary_index_map, ov.boundary_index_map)
    assert np.array_equal(gv.boundary_q_values.view(np.uint32), ov.boundary_q_values.view(np.uint32))


@pytest.mark.parametrize("workload", ["rocket512", "rocket1024", "building512"])
def test_gpu_voxelizer_bit_exact_at_full_size(workload):
    """The workloads' own geometry at full size: the plain form's finned rocket (about 20 k triangles) in the
    512x256x256 tunnel and, doubled, in the precision form's 1024x512x512, and the learned-closure form's block
    building in 512x128x256 with its base sunk into the floor. Flags, the boundary list and every q bit equal the
    oracle's."""
    from bench import rocket_workload
    from oracle import oracle as O

    from aerosim_amd import closure as C

    if workload.startswith("rocket"):
        scale = 1 if workload == "rocket512" else 2
        grid, H = (512 * scale, 256 * scale, 256 * scale), 256.0 * scale
        tris, _ = rocket_workload(*grid, float(scale))
        tunnel = A.Tunnel(float(grid[0]), float(grid[2]), H)
    else:
        case = C.BuildingCase()
        grid, H = (case.nx, case.height, case.lateral), float(case.height)
        tris, _ = case.mesh()
        tunnel = A.Tunnel(float(case.nx), float(case.lateral), H)
    ov = O.Voxels(tris, H, grid)
    gv = A.voxelize_mesh_to_lattice(tris, None, tunnel, grid)
    print(f"{workload}: {len(tris) // 9 if tris.ndim == 1 else len(tris)} triangles, solid {gv.solid_cell_count}, "
          f"boundary {gv.boundary_cell_count} ({gv.voxelization_time_ms:.2f} ms)")
    assert gv.boundary_cell_count > 1000
    assert np.array_equal(gv.cell_flags, ov.cell_flags)
    assert np.array_equal(gv.boundary_cell_indices, ov.boundary_cell_indices)
    assert np.array_equal(gv.boundary_q_values.view(np.uint32), ov.boundary_q_values.view(np.uint32))


@pytest.mark.parametrize("case", ["rocket", "naca", "building"])
def test_gpu_voxelizer_bit_exact_shapes(case):
    from oracle import oracle as O

    if case == "rocket":
        tris, H = rocket_case(128, 64, 64)
        grid = (128, 64, 64)
    elif case == "naca":
        grid = (96, 48, 24)
        tris = A.mesh_naca(0.12, 40.0, (-30.0, 0.0, 0.0), 8.0, -20.0, 20.0, 64)
        H = 48.0
    else:
        grid = (64, 32, 48)
        tris = A.mesh_box((-10.0, -17.0, -5.0), (0.0, 14.3, 5.0))
        H = 32.0
    ov = O.Voxels(tris, H, grid)
    gv = A.voxelize_mesh_to_lattice(tris, None, A.Tunnel(grid[0], grid[2], H), grid)
    print(f"{case}: solid {gv.solid_cell_count} boundary {gv.boundary_cell_count} ({gv.voxelization_time_ms:.2f} ms)")
    assert np.array_equal(gv.cell_flags, ov.cell_flags)
    assert np.array_equal(gv.boundary_cell_indices, ov.boundary_cell_indices)
    assert np.array_equal(gv.boundary_q_values.view(np.uint32), ov.boundary_q_values.view(np.uint32))
