"""GPU checks of the shifted-FP16 storage form and of the batched many-tunnel launch."""
import numpy as np
import pytest

import aerosim_amd as A
from tests.helpers import relerr, sphere_case

pytestmark = pytest.mark.gpu


def _sphere_ctx(storage, u=0.05, nu=0.004, cs=0.17, seed=1234):
    tris, H = sphere_case(64, 32, 32)
    ctx = A.Context(64, 32, 32, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    ctx.set_params(A.SolverParams(u_inlet=u, nu=nu, smagorinsky_cs=cs))
    ctx.reset(seed, 1e-4)
    return ctx, tris


def test_fp16_populations_close_to_fp32():
    c32, _ = _sphere_ctx(A.AERO_STORE_FP32)
    c16, _ = _sphere_ctx(A.AERO_STORE_FP16)
    c32.step(10, readout=True)
    c16.step(10, readout=True)
    flags = c32.read_voxels().cell_flags
    m = flags != A.CELL_SOLID
    f32 = c32.read_populations().reshape(19, -1)[:, m]
    f16 = c16.read_populations().reshape(19, -1)[:, m]
    err = relerr(f16, f32)
    print(f"fp16 vs fp32 populations after 10 steps: {err:.2e}")
    assert err < 2e-3


def test_fp16_cd_drift_c1():
    """Config-1 sphere, 1000 steps: FP16-store Cd against the FP64 oracle, tolerance 0.5 %."""
    from oracle import oracle as O

    u, nu = 0.05, 0.004
    c16, tris = _sphere_ctx(A.AERO_STORE_FP16, u=u, nu=nu)
    ov = O.Voxels(tris, 32.0, (64, 32, 32))
    orc = O.Solver(ov, "fp64", u_inlet=u, nu=nu, cs=0.17)
    orc.reset(1234, 1e-4)
    st = A.mesh_stats(tris)
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=32.0,
                tunnel_height=32.0, dx=1.0, u_ref=u)
    cg, co = [], []
    for s in range(1, 1001):
        c16.step(1)
        orc.step(1)
        if s > 960:
            cg.append(c16.read_forces(refs)["cd"])
            co.append(O.coefficients(orc.forces(), refs["frontal_area"], refs["char_length"], 32.0, 32.0, 1.0, u)["cd"])
    d = abs(np.mean(cg) - np.mean(co)) / abs(np.mean(co))
    print(f"C1 Cd fp16-store {np.mean(cg):.6f} vs fp64 oracle {np.mean(co):.6f}: |dCd| = {d * 100:.4f} %")
    assert d <= 0.005


def _naca_tunnel(nx, ny, nz, alpha, t, chord, Re, u=0.06, storage=A.AERO_STORE_FP32):
    tris = A.mesh_naca(t, chord, (-0.3 * nx, 0.0, 0.0), alpha, -nz, nz, 48)
    ctx = A.Context(nx, ny, nz, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
    ctx.set_params(A.SolverParams(u_inlet=u, nu=u * chord / Re, smagorinsky_cs=0.17))
    ctx.reset(5, 1e-4)
    return ctx


CASES = [(48, 24, 16, -4.0, 0.12, 12.0, 5e3), (64, 32, 24, 8.0, 0.18, 16.0, 2e4), (40, 24, 24, 16.0, 0.08, 10.0, 5e4),
         (56, 28, 20, 2.0, 0.15, 14.0, 1e4)]


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
@pytest.mark.parametrize("shape", [(256, 1, 512), (128, 4, 256), (512, 2, 1024), (256, 2, 0)])
def test_batch_matches_single_tunnels(shape, storage):
    singles = [_naca_tunnel(*c, storage=storage) for c in CASES]
    batched = [_naca_tunnel(*c, storage=storage) for c in CASES]
    for c in singles:
        c.step(21)
    b = A.Batch(batched)
    b.set_shape(*shape)
    b.step(21)
    for c in batched:
        assert c.step_index() == 21
    for cs, cb in zip(singles, batched):
        cs.step(1, readout=True)
        cb.step(1, readout=True)
        fs, fb = cs.read_populations(), cb.read_populations()
        assert np.array_equal(fs, fb)
        Fs, Fb = cs.read_forces(), cb.read_forces()
        assert Fs["fx"] == Fb["fx"] and Fs["fy"] == Fb["fy"]
    b.close()


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
@pytest.mark.parametrize("shape", [(256, 1, 512), (512, 2, 1024), (256, 2, 0)])
def test_batch_matches_single_tunnels_when_the_anchor_fires(shape, storage):
    """Populations raised by 0.2 % make the mass anchor fire. Every form sums the densities exactly (each is a
    float in [0.5, 1.8], so a multiple of 2^-24, and the sums stay far below 2^29), so the batched tunnels apply
    the single runs' anchor factor and keep their populations to the last bit."""
    singles = [_naca_tunnel(*c, storage=storage) for c in CASES]
    batched = [_naca_tunnel(*c, storage=storage) for c in CASES]
    for cs, cb in zip(singles, batched):
        cs.step(5, readout=True)
        cb.step(5, readout=True)
        f = cs.read_populations()
        assert np.array_equal(f, cb.read_populations())
        g = f * np.float32(1.002)
        cs.write_populations(g)
        cb.write_populations(g)
    before = [c.anchor_count() for c in singles]
    for c in singles:
        c.step(6)
    b = A.Batch(batched)
    b.set_shape(*shape)
    b.step(6)
    for cs, cb, a0 in zip(singles, batched, before):
        assert cs.anchor_count() > a0, "the raised density must trigger the anchor"
        assert cb.anchor_count() == cs.anchor_count()
        cs.step(1, readout=True)
        cb.step(1, readout=True)
        assert np.array_equal(cs.read_populations(), cb.read_populations())
        Fs, Fb = cs.read_forces(), cb.read_forces()
        assert Fs["fx"] == Fb["fx"] and Fs["fy"] == Fb["fy"]
    b.close()


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_batch_shape_switches_keep_single_tunnel_results(storage):
    """A batch that changes its shape between calls (one-launch grids, per-kind launches, other block sizes) and
    steps odd and even counts: the tables, the work counters and the density totals are rebuilt or re-armed at
    every switch, and the tunnels still match their single runs bit for bit."""
    singles = [_naca_tunnel(*c, storage=storage) for c in CASES]
    batched = [_naca_tunnel(*c, storage=storage) for c in CASES]
    b = A.Batch(batched)
    total = 0
    for shape, n in (((256, 1, 512), 3), ((256, 2, 0), 2), ((512, 1, 1024), 5), ((128, 2, 256), 1),
                     ((512, 1, 1024), 4), ((256, 1, 0), 3)):
        b.set_shape(*shape)
        b.step(n)
        total += n
    for c in singles:
        c.step(total)
    for cs, cb in zip(singles, batched):
        assert cb.step_index() == total
        cs.step(1, readout=True)
        cb.step(1, readout=True)
        assert np.array_equal(cs.read_populations(), cb.read_populations())
        Fs, Fb = cs.read_forces(), cb.read_forces()
        assert Fs["fx"] == Fb["fx"] and Fs["fy"] == Fb["fy"]
    b.close()


def test_batch_tuner_returns_a_valid_shape():
    ctxs = [_naca_tunnel(*c) for c in CASES]
    b = A.Batch(ctxs)
    r = b.tune(4)
    print("tuned shape", r)
    assert r["block"] in (128, 256, 512) and r["cells_per_thread"] in (1, 2, 4) and r["grid"] >= 0
    assert r["ms_per_step"] > 0
    with pytest.raises(A.AeroError):
        b.set_shape(256, 1, 100)  # the one-launch grid must be a multiple of 8
    b.close()


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_cells_per_thread_variants_identical(storage):
    """The 1/2/4 cells-per-thread step kernels give bit-identical populations."""
    for steps in (2, 12):
        outs = []
        for v in (1, 2, 4):
            c, _ = _sphere_ctx(storage)
            c.set_cells_per_thread(v)
            c.step(steps, readout=True)
            outs.append(c.read_populations().reshape(19, -1))
            c.close()
        for v, o in zip((2, 4), outs[1:]):
            bad = np.nonzero(o != outs[0])
            if len(bad[0]):
                cells = np.unique(bad[1])
                x, y, z = cells % 64, (cells // 64) % 32, cells // (64 * 32)
                print(f"steps {steps} V {v}: {len(bad[0])} values differ in {len(cells)} cells, max |d| "
                      f"{np.abs(o - outs[0]).max():.3e}, x {np.unique(x)[:12]}, y {np.unique(y)[:12]}, "
                      f"z {np.unique(z)[:12]}")
        assert all(np.array_equal(outs[0], o) for o in outs[1:]), f"after {steps} steps"


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_anchor_factor_identical_across_cells_per_thread(storage):
    """With the anchor firing, the per-cell step (cells per thread 0) and the split step at 1, 2 and 4 cells per
    thread reduce the densities in different orders and still give bit-identical populations."""
    base, _ = _sphere_ctx(storage)
    base.step(3, readout=True)
    g = base.read_populations() * np.float32(1.002)
    base.close()
    outs = []
    for v in (0, 1, 2, 4):
        c, _ = _sphere_ctx(storage)
        c.set_cells_per_thread(v)
        c.write_populations(g)
        a0 = c.anchor_count()
        c.step(5, readout=True)
        assert c.anchor_count() > a0
        outs.append((c.anchor_count(), c.read_populations()))
        c.close()
    for v, (n, f) in zip((1, 2, 4), outs[1:]):
        assert n == outs[0][0], f"cells per thread {v}"
        assert np.array_equal(f, outs[0][1]), f"cells per thread {v}"


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_split_step_matches_per_cell_step(storage):
    """The bulk + special-cell launches, plain and readout steps alike, give the populations of the
    one-cell-per-thread kernel (cells per thread 0)."""
    a, _ = _sphere_ctx(storage)
    b, _ = _sphere_ctx(storage)
    a.set_cells_per_thread(0)
    a.step(12, readout=True)
    for _ in range(12):
        b.step(1, readout=True)  # every step a split readout step
    pa, pb = a.read_populations(), b.read_populations()
    print("max |split - per-cell|", float(np.abs(pa - pb).max()))
    assert np.array_equal(pa, pb)
    a.close()
    b.close()


def _plates_ctx(storage):
    """The sphere plus one-cell plates in the inlet (x = 0) and outlet (x = nx-1) planes. At a plate's corners a
    cell at x = 0 (nx-1) is a boundary cell through its (0, +-1, +-1) link while its x = 1 (nx-2) neighbour in
    the same row is plain fluid: the bulk test must then leave the x = 1 (nx-2) group out, since that group
    also writes the row's inlet (outlet) cell."""
    tris, H = sphere_case(64, 32, 32)
    plates = [A.mesh_box((-33.0, -4.0, -3.0), (-31.2, 4.0, 5.0)), A.mesh_box((31.2, -6.0, -5.0), (33.0, 2.0, 3.0))]
    ctx = A.Context(64, 32, 32, storage=storage)
    ctx.set_mesh(np.concatenate([tris] + plates), A.Tunnel(64.0, 32.0, H))
    ctx.set_params(A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17))
    ctx.reset(1234, 1e-4)
    return ctx


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_split_step_matches_per_cell_step_with_inlet_outlet_walls(storage):
    c = _plates_ctx(storage)
    fl = c.read_voxels().cell_flags.reshape(32, 32, 64)  # [z, y, x]
    interior = np.zeros((32, 32), dtype=bool)
    interior[1:-1, 1:-1] = True
    for xe, xn in ((0, 1), (63, 62)):
        case = interior & (fl[:, :, xe] != A.CELL_FLUID) & (fl[:, :, xn] == A.CELL_FLUID)
        print(f"x = {xe}: {int(case.sum())} interior rows with a non-fluid end cell beside plain fluid")
        assert case.any()
    c.close()
    for v in (1, 2, 4):
        a, b = _plates_ctx(storage), _plates_ctx(storage)
        a.set_cells_per_thread(v)
        a.step(13, readout=True)  # twelve split steps, then one split readout step
        b.set_cells_per_thread(0)
        b.step(13, readout=True)  # every step through the per-cell kernel
        pa, pb = a.read_populations(), b.read_populations()
        a.close()
        b.close()
        print(f"V {v}: max |split - per-cell| {float(np.abs(pa - pb).max()):.3e}")
        assert np.array_equal(pa, pb)


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_wide_cross_section(storage):
    """An 8192 x 4096 x 3 tunnel (one bulk layer, z = 1; 19 N stays below 2^31). FP32 steps with flat 64-bit
    addresses. For FP16 the bulk kernel's 32-bit buffer offsets (2 x 19 nx (ny + 1) elements) would pass 2^31
    bytes, and the context steps with the per-cell kernel instead. Either way the flow stays finite with its
    mass."""
    assert 2 * 19 * 8192 * 4097 * 2 > 2 ** 31 and 19 * 8192 * 4096 * 3 < 2 ** 31
    import aerosim_amd as A

    ctx = A.Context(8192, 4096, 3, storage=A.AERO_STORE_FP16 if storage == "fp16" else A.AERO_STORE_FP32)
    ctx.set_mesh(A.mesh_box((-100.0, -100.0, -0.6), (100.0, 100.0, 0.6)), A.Tunnel(8192.0, 3.0, 4096.0))
    ctx.set_params(A.SolverParams(u_inlet=0.05, nu=0.01, smagorinsky_cs=0.17))
    ctx.reset(3, 1e-4)
    ctx.step(3)
    ctx.step(1, readout=True)
    s = ctx.mass_stats()
    ctx.close()
    print(f"8192x4096x3 {storage}: mean rho {s['mean_rho']:.6f}, max |u| {s['max_speed']:.4f}")
    assert not s["has_nan"] and abs(s["mean_rho"] - 1.0) < 1e-3
