"""GPU parity of the derived-field readouts against the oracle (oracle/fields_oracle.py), bit for bit, on
the macro fields of a developed sphere wake (read back through aero_read_macro, which applies the same
mass-anchor factor as the kernels), and the reference's phase5 readout tests (src/validation/
phase5.test.ts:13-177) through the GPU path."""
import numpy as np
import pytest

import aerosim_amd as A
from oracle import fields_oracle as F
from tests.helpers import sphere_case

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 32, 32
U = 0.06


@pytest.fixture(scope="module")
def flow():
    tris, _ = sphere_case(NX, NY, NZ)
    ctx = A.Context(NX, NY, NZ)
    ctx.set_mesh(tris, A.Tunnel(float(NX), float(NZ), float(NY)))
    ctx.set_params(A.SolverParams(u_inlet=U, nu=0.004, smagorinsky_cs=0.17))
    ctx.reset(5, 1e-3)
    ctx.step(300, readout=True)
    ux, uy, uz, rho = ctx.read_macro()
    flags = ctx.read_voxels(float(NY)).cell_flags
    print(f"sphere wake: anchor fired {ctx.anchor_count()} times, max |u| {np.sqrt(ux**2 + uy**2 + uz**2).max():.4f}")
    yield ctx, (ux, uy, uz, rho), np.asarray(flags, dtype=np.uint8)
    ctx.close()


def _diff(a, b):
    bad = np.argwhere(a != b)
    return f"{len(bad)} mismatches, first {bad[:3].tolist()}: {a[tuple(bad[0])] if len(bad) else ''} vs " \
           f"{b[tuple(bad[0])] if len(bad) else ''}"


@pytest.mark.parametrize("rules", ["cpu", "wgsl"])
def test_slice_rgba_matches_oracle(flow, rules):
    ctx, (ux, uy, uz, rho), flags = flow
    ref = F.slice_rgba_cpu if rules == "cpu" else F.slice_rgba_wgsl
    for axis, an in ((0, "z"), (1, "y"), (2, "x")):
        for field, fn in ((0, "velocity"), (1, "vorticity"), (2, "cp")):
            for frac in (0.0, 0.31, 0.5, 1.0):
                w, h, img = ctx.extract_slice_rgba(an, frac, fn, rules, U)
                want = ref(NX, NY, NZ, ux, uy, uz, rho, flags, axis, frac, field, U)
                assert img.shape == want.shape == (h, w, 4)
                assert np.array_equal(img, want), f"{rules} {an} {fn} {frac}: {_diff(img, want)}"
    # the slice is not blank: the wake spans several colours and the sphere is transparent
    w, h, img = ctx.extract_slice_rgba("z", 0.5, "vorticity", rules, U)
    assert len(np.unique(img.reshape(-1, 4), axis=0)) > 20 and (img[..., 3] == 0).any()


def test_qcriterion_matches_oracle(flow):
    ctx, (ux, uy, uz, rho), flags = flow
    got = ctx.compute_q_criterion(dx=0.05)
    want = F.q_criterion(NX, NY, NZ, 0.05, ux, uy, uz, flags)
    for k in ("q_values", "omega_mag", "omega_vec"):
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), f"{k}: {_diff(got[k], want[k])}"
    assert got["max_q"] == want["max_q"] and got["max_omega"] == want["max_omega"]
    print(f"max Q {got['max_q']:.6e}, max |omega| {got['max_omega']:.6e}")
    assert got["max_q"] > 1e-6


def test_qcriterion_standalone_on_supplied_fields():
    rng = np.random.default_rng(8)
    for nx, ny, nz in ((16, 16, 16), (1, 5, 7), (33, 2, 1), (9, 11, 13)):
        N = nx * ny * nz
        ux, uy, uz = (rng.standard_normal(N).astype(np.float32) for _ in range(3))
        flags = (rng.random(N) < 0.2).astype(np.uint8)
        for fl in (None, flags):
            got = A.compute_q_criterion(nx, ny, nz, 0.1, ux, uy, uz, fl)
            want = F.q_criterion(nx, ny, nz, 0.1, ux, uy, uz, fl)
            for k in ("q_values", "omega_mag", "omega_vec"):
                assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), (nx, ny, nz, k)
            assert got["max_q"] == want["max_q"] and got["max_omega"] == want["max_omega"]


def test_reference_phase5_rotation_and_shear_on_gpu():
    nx = ny = nz = 16
    dx, om0 = 0.1, 5.0
    c = (nx - 1) / 2
    iz, iy, ix = (a.reshape(-1) for a in np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij"))
    mid = nx // 2 + nx * (ny // 2 + ny * (nz // 2))
    r = A.compute_q_criterion(nx, ny, nz, dx, -om0 * (iy - c) * dx, om0 * (ix - c) * dx, 0 * ix)
    assert 20.0 < r["q_values"][mid] <= 25.5 and round(float(r["omega_mag"][mid])) == 10
    r = A.compute_q_criterion(nx, ny, nz, dx, 4.0 * (iy * dx), 0 * ix, 0 * ix)
    assert abs(r["q_values"][mid]) < 1e-4


def test_probe_matches_oracle(flow):
    ctx, (ux, uy, uz, rho), flags = flow
    rng = np.random.default_rng(9)
    origin, dx = (-3.2, -1.6, -1.6), 0.1
    pts = np.column_stack([rng.uniform(-3.5, 3.5, 400), rng.uniform(-1.9, 1.9, 400), rng.uniform(-1.9, 1.9, 400)])
    # cell centres, the sphere's cells, grid corners and faces
    solid_cells = np.flatnonzero(flags == 1)[:40]
    sx, sy, sz = solid_cells % NX, (solid_cells // NX) % NY, solid_cells // (NX * NY)
    extra = np.column_stack([origin[0] + (sx + 0.5) * dx, origin[1] + (sy + 0.5) * dx, origin[2] + (sz + 0.5) * dx])
    corners = np.array([[origin[0], origin[1], origin[2]], [origin[0] + NX * dx, origin[1] + NY * dx, origin[2] + NZ * dx],
                        [origin[0] + 0.5 * dx, 0.0, 0.0]])
    pts = np.vstack([pts, extra, corners])
    surf = (rng.random(len(pts)) < 0.1).astype(np.uint8)
    got = ctx.probe(pts, origin, dx, 30.0, U, 1.204, surf)
    want = F.probe(NX, NY, NZ, ux, uy, uz, rho, flags, pts, origin, dx, 30.0, U, 1.204, surf)
    assert np.array_equal(got, want), _diff(got, want)
    assert (got[:, 12] == 1).sum() >= 40 and (got[:, 12] == 0).sum() > 300


def _cube_voxels(grid, scale, tunnel):
    box = A.mesh_box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    world = np.diag([scale, scale, scale, 1.0]).reshape(-1)
    return A.voxelize_mesh_to_lattice(box, world, tunnel, grid)


def test_reference_phase5_stagnation_and_wake_cp():
    """phase5.test.ts:119-177: the preset cube scaled to 0.45 m in a 3 x 1.5 x 1.5 m tunnel at 40 x 20 x 20,
    u = 0.08, nu = 0.004, Cs = 0.12, 90 steps. On the mid row, the fluid cell in front of the cube has
    Cp in (0.65, 1.35) and the fluid cell behind it Cp < 0, with Cp = (rho - 1) / 3 / (u^2 / 2)."""
    from aerosim_amd import presets as P

    world = np.diag([0.45, 0.45, 0.45, 1.0]).reshape(-1)
    vox = A.voxelize_mesh_to_lattice(P.create_preset_geometry("cube"), world, A.Tunnel(3.0, 1.5, 1.5), (40, 20, 20))
    u = 0.08
    s = A.LBMSolver(vox, u_inlet=u, nu_lattice=0.004, smagorinsky_cs=0.12)
    s.step(90)
    nx, ny, nz = 40, 20, 20
    my, mz = ny // 2, nz // 2
    row = vox.cell_flags.reshape(nz, ny, nx)[mz, my]
    solid = np.nonzero(row == A.CELL_SOLID)[0]
    windward, leeward = int(solid[0]) - 1, int(solid[-1]) + 1
    assert windward > 0 and leeward > windward
    rho = s.rho
    q = 0.5 * u * u
    cp = lambda ix: ((rho[ix + nx * (my + ny * mz)] - 1.0) / 3.0) / q  # noqa: E731
    print(f"cube x in [{solid[0]}, {solid[-1]}]: stagnation Cp {cp(windward):.3f}, wake Cp {cp(leeward):.3f}")
    assert 0.65 < cp(windward) < 1.35
    assert cp(leeward) < 0.0


def test_reference_phase5_wake_vortex_cores():
    """phase5.test.ts:83-127: a 0.4 m cube in a 3 x 1.5 x 1.5 m tunnel at 36 x 18 x 18, 80 steps."""
    vox = _cube_voxels((36, 18, 18), 0.4, A.Tunnel(3.0, 1.5, 1.5))
    s = A.LBMSolver(vox, u_inlet=0.085, nu_lattice=0.005, smagorinsky_cs=0.12)
    s.step(80)
    r = s.ctx.compute_q_criterion(vox.dx)
    print(f"cube wake: max Q {r['max_q']:.4e}, max |omega| {r['max_omega']:.4e}")
    assert r["max_q"] > 1e-4 and r["max_omega"] > 0.01
    img = s.extract_slice_rgba("z", 0.5, "vorticity")
    assert img["width"] == 36 and img["height"] == 18 and img["rgba"].size == 36 * 18 * 4
    s.close()


def test_reference_phase5_stagnation_and_wake_cp():
    """phase5.test.ts:130-177: a 0.45 m cube at 40 x 20 x 20, 90 steps: Cp in (0.65, 1.35) in front of the
    cube, Cp < 0 behind it, on the mid row."""
    vox = _cube_voxels((40, 20, 20), 0.45, A.Tunnel(3.0, 1.5, 1.5))
    u = 0.08
    s = A.LBMSolver(vox, u_inlet=u, nu_lattice=0.004, smagorinsky_cs=0.12)
    s.step(90)
    nx, ny, nz = 40, 20, 20
    flags = np.asarray(vox.cell_flags)
    midY, midZ = ny // 2, nz // 2
    row = [flags[ix + nx * (midY + ny * midZ)] for ix in range(nx)]
    solid = [ix for ix in range(nx) if row[ix] == 1]
    wind, lee = solid[0] - 1, solid[-1] + 1
    q = 0.5 * u * u
    rho = s.rho
    cps = [((1 / 3) * (float(rho[i + nx * (midY + ny * midZ)]) - 1.0)) / q for i in (wind, lee)]
    print(f"stagnation Cp {cps[0]:.4f}, wake Cp {cps[1]:.4f}")
    assert 0.65 < cps[0] < 1.35 and cps[1] < 0.0
    # the probe reads the same Cp at the stagnation cell's centre
    p = s.ctx.probe([[vox.origin[0] + (wind + 0.5) * vox.dx, vox.origin[1] + (midY + 0.5) * vox.dx,
                      vox.origin[2] + (midZ + 0.5) * vox.dx]], vox.origin, vox.dx, 25.0, u, 1.204)
    assert abs(p[0, 9] - cps[0]) < 1e-9
    s.close()


def test_node_engine_slice_matches_python_engine():
    """The JavaScript LBMEngine's stepAndExtractSlice returns the slice of the Python LBMEngine, byte for byte."""
    import json
    import os
    import shutil
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    node = shutil.which("node")
    if node is None or not os.path.exists(os.path.join(root, "aerosim_amd", "node", "aerolbm.node")):
        pytest.skip("node or the addon is absent")
    script = """
const m = require('./aerosim_amd/node/lbmEngine.js');
const ico = m.addon.meshIcosphere(2, 0.4, [0, 0, 0]);
const tunnel = {length: 3.0, width: 1.5, height: 1.5};
const vox = m.voxelizeMeshToLattice({position: ico}, null, tunnel, {nx: 40, ny: 20, nz: 20});
const e = new m.LBMEngine(0);
e.setVoxelGrid(vox, {uInletLattice: 0.07, nuLattice: 0.004, smagorinskyCs: 0.12, tunnel, fieldMode: 'vorticity'});
e.stepAndExtractSlice(12, false).then((r) => {
  const q = e.computeQCriterion();
  console.log(JSON.stringify({w: r.width, h: r.height, rgba: Array.from(r.rgba), maxQ: q.maxQ}));
  e.destroy();
});
"""
    p = subprocess.run([node, "-e", script], cwd=root, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    r = json.loads(p.stdout)
    vox = A.voxelize_mesh_to_lattice(A.mesh_icosphere(2, 0.4, (0, 0, 0)), None, A.Tunnel(3.0, 1.5, 1.5), (40, 20, 20))
    eng = A.LBMEngine()
    eng.update_config(u_inlet_lattice=0.07, nu_lattice=0.004, smagorinsky_cs=0.12, tunnel=A.Tunnel(3.0, 1.5, 1.5))
    eng.set_voxel_grid(vox)
    out = eng.step_and_extract_slice(12)
    assert (r["w"], r["h"]) == (out["width"], out["height"]) == (40, 20)
    assert np.array_equal(np.array(r["rgba"], dtype=np.uint8), out["rgba"])
    assert r["maxQ"] == eng.ctx.compute_q_criterion(vox.dx)["max_q"]
    eng.ctx.close()
