"""Pins of the oracle's domain-boundary rules (lbmCpuReference.ts:114-143, 315-456) by moments.

Every expected value is computed here from this file's own D3Q19 tables and from the oracle's macro fields,
never from the oracle's equilibrium: an equilibrium of density rho and velocity u has zeroth moment rho, first
moments rho u and second moments rho (u_a u_b + delta_ab / 3), which is what each rule must leave behind. The
case is a small tunnel (nz > 4, so side walls apply) with a block standing on the floor, so the floor holds
BOUNDARY cells that the wall rules must skip. The flow carries 1e-3 seeded noise, so every velocity component the
rules set or keep is non-zero and a slip (a wrong component zeroed, a wrong cell class walled, post-wall values
read by the outlet, the anchor applied to solids) moves a checked moment by far more than the tolerance.
"""
import numpy as np
import pytest

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C = np.stack([CX, CY, CZ])  # 3 x 19
SOLID, BOUNDARY, FLUID = 1, 2, 0
NX, NY, NZ = 20, 10, 8
TOL = 1e-13


def _box(lo, hi):
    """Triangles (12 x 9 floats) of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float32)
    faces = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (3, 6, 2), (3, 7, 6),
             (0, 4, 7), (0, 7, 3), (1, 2, 6), (1, 6, 5)]
    return np.concatenate([v[list(f)].reshape(-1) for f in faces])


@pytest.fixture(scope="module")
def case(oracle_lib):
    O = oracle_lib
    # block on the floor: cells ix 7..10, iy 0..2, iz 3..4 (cell size 1, lattice centred on the origin)
    vox = O.Voxels(_box((-2.9, -6.0, -1.4), (0.9, -2.1, 1.4)), float(NY), (NX, NY, NZ))
    return O, vox


def _solver(O, vox, noslip=False, anchor=True):
    s = O.Solver(vox, "fp64", u_inlet=0.05, nu=0.01, cs=0.12, ground_noslip=noslip)
    s.set_anchor(anchor)
    s.reset(21, 1e-3)
    return s


def _fields(s):
    f = s.populations().reshape(19, NZ, NY, NX)
    ux, uy, uz, rho = (a.reshape(NZ, NY, NX) for a in s.macro())
    return f, np.stack([ux, uy, uz]), rho


def _moments(f):
    """Zeroth, first and second moments of populations f (19 x ...)."""
    m0 = f.sum(axis=0)
    m1 = np.tensordot(C, f, axes=(1, 0))
    m2 = np.einsum("ak,bk,k...->ab...", C, C, f)
    return m0, m1, m2


def _assert_equilibrium(f, rho, u, where):
    """f (19 x n) is the D3Q19 equilibrium of rho (n) and u (3 x n)."""
    m0, m1, m2 = _moments(f)
    assert np.max(np.abs(m0 - rho)) < TOL, where
    assert np.max(np.abs(m1 - rho * u)) < TOL, where
    eq2 = rho * (u[:, None] * u[None, :] + np.eye(3)[:, :, None] / 3.0)
    assert np.max(np.abs(m2 - eq2)) < TOL, where


def test_block_puts_boundary_cells_on_the_floor(case):
    O, vox = case
    fl = np.asarray(vox.cell_flags).reshape(NZ, NY, NX)
    assert np.all(fl[3:5, 0:3, 7:11] == SOLID)
    floor = fl[:, 0, 1:NX - 1]
    assert (floor == BOUNDARY).sum() >= 8 and (floor == FLUID).sum() > 50


def test_free_slip_floor_ceiling_and_side_walls(case):
    O, vox = case
    s = _solver(O, vox)
    s.step(30)
    assert s.anchor_count() > 0  # the small tunnel's mass drifts: the walls are checked after anchored steps
    f, u, rho = _fields(s)
    fl = np.asarray(vox.cell_flags).reshape(NZ, NY, NX)
    xs = slice(1, NX - 1)
    for iy, name in ((0, "floor"), (NY - 1, "ceiling")):
        m = fl[:, iy, xs] == FLUID
        fw, uw, rw = f[:, :, iy, xs][:, m], u[:, :, iy, xs][:, m], rho[:, iy, xs][m]
        assert np.all(uw[1] == 0.0), name  # the macro field records the wall's u_y = 0
        _assert_equilibrium(fw, rw, uw, name)
        assert np.max(np.abs(uw[0])) > 1e-3 and np.max(np.abs(uw[2])) > 1e-5, name  # u_x, u_z kept, non-zero
    for iz, name in ((0, "side z = 0"), (NZ - 1, "side z = nz-1")):
        m = fl[iz, :, xs] == FLUID
        fw, uw, rw = f[:, iz, :, xs][:, m], u[:, iz, :, xs][:, m], rho[iz, :, xs][m]
        assert np.all(uw[2] == 0.0), name
        _assert_equilibrium(fw, rw, uw, name)
        assert np.max(np.abs(uw[1][np.nonzero(m)[0] != 0])) > 1e-5, name  # u_y kept off the floor
    # floor-side corners: both wall components vanish
    for iz in (0, NZ - 1):
        m1 = _moments(f[:, iz, 0, xs])[1]
        assert np.max(np.abs(m1[1])) < TOL and np.max(np.abs(m1[2])) < TOL


def test_walls_skip_boundary_cells_and_no_slip_floor_is_at_rest(case):
    """One step from the same state with the free-slip and the no-slip floor: the floor's BOUNDARY cells come out
    identical (no wall rule touches them, lbmCpuReference.ts:380,394), its FLUID cells differ, and under no-slip
    they hold the rest equilibrium w_k rho (sum c f = 0). The anchor is off: its factor would differ between the
    two runs."""
    O, vox = case
    a = _solver(O, vox, anchor=False)
    b = _solver(O, vox, noslip=True, anchor=False)
    a.step(30)
    b.step(30)
    b.set_populations(a.populations())
    a.step(1)
    b.step(1)
    assert a.anchor_count() == 0 and b.anchor_count() == 0
    fa, ua, ra = _fields(a)
    fb, ub, rb = _fields(b)
    fl = np.asarray(vox.cell_flags).reshape(NZ, NY, NX)
    xs = slice(1, NX - 1)
    bnd = fl[:, 0, xs] == BOUNDARY
    flu = fl[:, 0, xs] == FLUID
    assert np.array_equal(fa[:, :, 0, xs][:, bnd], fb[:, :, 0, xs][:, bnd])
    # a BOUNDARY floor cell is left as the collision wrote it: its u_y is not the wall's zero
    assert np.max(np.abs(ua[1][:, 0, xs][bnd])) > 1e-6
    assert np.max(np.abs(fa[:, :, 0, xs][:, flu] - fb[:, :, 0, xs][:, flu])) > 1e-5
    # no-slip: f = w rho exactly on the floor's fluid cells away from the side walls (the corner cells then get
    # the side rule, which keeps u = 0 there too)
    fw, rw = fb[:, :, 0, xs][:, flu], rb[:, 0, xs][flu]
    assert np.max(np.abs(fw - W[:, None] * rw[None, :])) < TOL
    assert np.all(ub[:, :, 0, xs][:, flu] == 0.0)


def _profile(kind, iy, u_in, z0=0.5, zref=6.0):
    if kind == "power":  # the reference's ground profile, lbmCpuReference.ts:122-125, 335-338
        return u_in * np.minimum(1.15, 0.35 + 0.8 * ((iy + 0.5) / NY) ** 0.3)
    # the log law of the boundary-layer workload: u_in ln(z / z0) / ln(z_ref / z0) at z = iy + 1/2, not below 0
    return u_in * np.maximum(0.0, np.log((iy + 0.5) / z0) / np.log(zref / z0))


@pytest.mark.parametrize("kind", ["power", "log"])
def test_ground_inflow_profile_init_and_inlet(case, kind):
    O, vox = case
    s = O.Solver(vox, "fp64", u_inlet=0.05, nu=0.01, cs=0.12, ground_noslip=True, profile=2 if kind == "log" else 0,
                 z0=0.5, zref=6.0)
    s.set_anchor(False)  # rho_in = 0.85 rho(ix = 1) + 0.15 holds before an anchor rescales both
    s.reset(5, 0.0)
    f, u, rho = _fields(s)
    fl = np.asarray(vox.cell_flags).reshape(NZ, NY, NX)
    iy = np.arange(NY)[None, :, None] * np.ones((NZ, NY, NX))
    expect_ux = np.where(iy == 0, 0.0, _profile(kind, iy, 0.05))
    expect_ux = np.where(fl == SOLID, 0.0, expect_ux)
    m0, m1, _ = _moments(f)
    assert np.max(np.abs(m0 - 1.0)) < TOL
    assert np.max(np.abs(m1[0] - expect_ux)) < TOL and np.max(np.abs(m1[1:])) < TOL
    for steps in (7, 150):  # inside and past the 120-step inlet perturbation
        s.step(steps - s.step_index)
        step_of_rule = s.step_index - 1  # the inlet rule of the last step saw this step index
        assert s.anchor_count() == 0
        f, u, rho = _fields(s)
        fi, ri, r1 = f[:, :, :, 0], rho[:, :, 0], rho[:, :, 1]
        assert np.max(np.abs(ri - (0.85 * r1 + 0.15))) < TOL
        rows = np.arange(NY)[None, :] * np.ones((NZ, NY))
        uy = 0.015 * 0.05 * np.sin(step_of_rule * 0.15 + rows / NY * 2 * np.pi) if step_of_rule < 120 else 0 * rows
        uin = np.stack([_profile(kind, rows, 0.05), uy, 0 * rows])
        _assert_equilibrium(fi.reshape(19, -1), ri.reshape(-1), uin.reshape(3, -1), f"inlet after {steps} steps")


def test_outlet_reads_the_pre_wall_values_of_its_neighbour(case):
    """The outlet (lbmCpuReference.ts:354-371) runs before the wall rules, so on the floor it takes u_y of its
    ix = nx-2 neighbour as that cell's collision left it (not the wall's 0). That u_y is recomputed here from the
    pre-step populations: pulled with the clamp at the domain faces, rho clamped to [0.5, 1.8], u = m / rho."""
    O, vox = case
    s = _solver(O, vox, anchor=False)
    s.step(30)
    f0 = s.populations().reshape(19, NZ, NY, NX)
    s.step(1)
    f1, u1, r1 = _fields(s)
    x = NX - 2
    for iz in range(1, NZ - 1):
        pull = np.array([f0[k, min(max(iz - CZ[k], 0), NZ - 1), min(max(0 - CY[k], 0), NY - 1), x - CX[k]]
                         for k in range(19)])
        r = min(1.8, max(0.5, pull.sum()))
        uy_pre = (CY * pull).sum() / r
        assert abs(u1[1, iz, 0, x]) == 0.0  # the wall rule zeroed u_y at ix = nx-2
        assert abs(uy_pre) > 1e-6
        assert abs(u1[1, iz, 0, NX - 1] - uy_pre) < 1e-14
        # f_out = 0.75 f_prev + 0.25 feq(rho_out, u_prev): its mass is 0.75 rho_prev + 0.25 rho_out
        rho_out = 0.85 * r1[iz, 0, x] + 0.15
        assert abs(r1[iz, 0, NX - 1] - rho_out) < TOL
        assert abs(f1[:, iz, 0, NX - 1].sum() - (0.75 * r1[iz, 0, x] + 0.25 * rho_out)) < TOL


def test_mass_anchor_scales_non_solid_cells_by_one_factor(case):
    """One step from a given state with and without the anchor (lbmCpuReference.ts:431-455). When the step's mean
    non-solid density is off by 1e-3, the anchored result equals the unanchored one times 1/mean on every non-solid
    population and density (one common factor), solid cells are untouched, and the mean is 1 afterwards. When it
    is off by less than the 1e-4 threshold, the two results are identical. The states are the same populations
    scaled on the non-solid cells, the scale chosen from the unanchored step's mean."""
    O, vox = case
    base = _solver(O, vox)
    base.step(12)
    f = base.populations().reshape(19, -1)
    nonsolid = np.asarray(vox.cell_flags) != SOLID

    def run(scale):
        g = f.copy()
        g[:, nonsolid] *= scale
        on, off = _solver(O, vox), _solver(O, vox, anchor=False)
        for s in (on, off):
            s.step(12)
            s.set_populations(g.reshape(-1))
        n0 = on.anchor_count()
        for s in (on, off):
            s.step(1)
        return on.anchor_count() - n0, on, off

    _, _, probe = run(1.0)
    m1 = probe.macro()[3][nonsolid].mean()
    for scale, fires in ((1.001 / m1, True), (1.0 / m1, False)):
        fired, on, off = run(scale)
        f_on, f_off = on.populations().reshape(19, -1), off.populations().reshape(19, -1)
        r_on, r_off = on.macro()[3], off.macro()[3]
        mean_off = r_off[nonsolid].sum() / nonsolid.sum()
        assert fired == (1 if fires else 0), (scale, mean_off)
        if fires:
            assert abs(mean_off - 1.0) > 1e-4
            # one factor on every non-solid population and density (each product rounds once: 2^-53 apart at most),
            # and that factor is 1/mean (the mean summed here in another order than the oracle's loop)
            ratio = np.concatenate([(f_on[:, nonsolid] / f_off[:, nonsolid]).ravel(), r_on[nonsolid] / r_off[nonsolid]])
            assert (ratio.max() - ratio.min()) / ratio.min() < 4e-16
            assert abs(np.median(ratio) * mean_off - 1.0) < 1e-13
            assert abs(r_on[nonsolid].mean() - 1.0) < 1e-12
        else:
            assert abs(mean_off - 1.0) <= 1e-4
            assert np.array_equal(f_on, f_off)
        assert np.array_equal(f_on[:, ~nonsolid], f_off[:, ~nonsolid])
