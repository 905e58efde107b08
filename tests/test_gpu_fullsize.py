"""Size-independent properties at the benchmark's full size, the 512x256x256 finned-rocket tunnel (33.5 M
cells), and at the precision form's 1024x512x512: the split step (bulk kernel, the listed cells beside it,
graph-replayed pairs) gives the populations of the one-cell-per-thread kernel bit for bit, a second run from the
same seed repeats the first bit for bit, and the mean density stays at 1."""
import numpy as np
import pytest

import aerosim_amd as A
from bench import rocket_workload

pytestmark = pytest.mark.gpu

NX, NY, NZ = 512, 256, 256


def _ctx(storage):
    tris, D = rocket_workload(NX, NY, NZ, 1.0)
    ctx = A.Context(NX, NY, NZ, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(NX), float(NZ), float(NY)))
    ctx.set_params(A.SolverParams(u_inlet=0.08, nu=0.08 * D / 2e4, smagorinsky_cs=0.17))
    ctx.reset(7, 1e-4)
    return ctx


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_fullsize_split_step_matches_per_cell_step(storage):
    steps = 21
    a = _ctx(storage)
    a.step(steps, readout=True)  # 20 plain steps (ten replays of a captured pair), then a readout step
    info = a.split_info()
    pa = a.read_populations()
    s = a.mass_stats()
    a.close()
    b = _ctx(storage)
    b.set_cells_per_thread(0)
    b.step(steps, readout=True)  # every step through the per-cell kernel
    pb = b.read_populations()
    b.close()
    diff = int(np.count_nonzero(pa != pb))
    print(f"512x256x256 storage {storage}: {info}; {diff} of {pa.size} populations differ; mean rho "
          f"{s['mean_rho']:.7f}, max |u| {s['max_speed']:.4f}")
    assert info["bulk_cells"] > 30_000_000
    assert diff == 0
    assert not s["has_nan"] and abs(s["mean_rho"] - 1.0) < 1e-3


def _digest(a):
    import xxhash

    return xxhash.xxh3_128_hexdigest(np.ascontiguousarray(a).view(np.uint8))


def test_precision_form_size_split_step_matches_per_cell_step():
    """The precision form's own size, 1024 x 512 x 512 in shifted-FP16 storage (268 M cells): 19 N = 5.1 G population
    slots, past 2^32, where any 32-bit slot index would wrap. After 9 steps the split step and the one-cell-per-thread
    kernel give the same 5.1 G populations (compared through an xxh3-128 digest of each 20 GB read-back), the same
    macro fields and the same forces, and the mean density stays at 1."""
    nx, ny, nz = 1024, 512, 512
    tris, D = rocket_workload(nx, ny, nz, 2.0)
    out = []
    for cpt in (None, 0):
        c = A.Context(nx, ny, nz, storage=A.AERO_STORE_FP16)
        c.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
        c.set_params(A.SolverParams(u_inlet=0.08, nu=0.08 * D / 2e4, smagorinsky_cs=0.17))
        c.reset(7, 1e-4)
        if cpt is not None:
            c.set_cells_per_thread(cpt)
        c.step(9, readout=True)  # 8 plain steps (four replays of a captured pair), then a readout step
        rec = {"pops": _digest(c.read_populations()), "macro": [_digest(m) for m in c.read_macro()],
               "fx": c.read_forces()["fx"], "stats": c.mass_stats(), "info": c.split_info()}
        c.close()
        out.append(rec)
    a, b = out
    print(f"1024x512x512 fp16: {a['info']}; populations {a['pops']} / {b['pops']}; mean rho "
          f"{a['stats']['mean_rho']:.7f}, max |u| {a['stats']['max_speed']:.4f}")
    assert a["info"]["bulk_cells"] > 260_000_000
    assert a["pops"] == b["pops"] and a["macro"] == b["macro"] and a["fx"] == b["fx"]
    assert not a["stats"]["has_nan"] and abs(a["stats"]["mean_rho"] - 1.0) < 1e-3


def test_fullsize_runs_repeat_bit_for_bit():
    runs = []
    for _ in range(2):
        c = _ctx(A.AERO_STORE_FP32)
        c.step(41, readout=True)
        runs.append(c.read_populations())
        c.close()
    assert np.array_equal(runs[0], runs[1])


def _cd_window(storage, seed, steps=5000, window=2000):
    tris, D = rocket_workload(NX, NY, NZ, 1.0)
    ctx = A.Context(NX, NY, NZ, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(NX), float(NZ), float(NY)))
    ctx.set_params(A.SolverParams(u_inlet=0.08, nu=0.08 * D / 2e4, smagorinsky_cs=0.17))
    ctx.reset(seed, 1e-4)
    st = A.mesh_stats(tris)
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=float(NZ),
                tunnel_height=float(NY), dx=1.0, u_ref=0.08)
    ctx.step(steps - window)
    cd = np.zeros(window)
    for i in range(window):
        ctx.step(1, readout=i == window - 1)
        cd[i] = ctx.read_forces(refs)["cd"]
    s = ctx.mass_stats()
    ctx.close()
    assert not s["has_nan"] and abs(s["mean_rho"] - 1.0) < 1e-3
    return float(cd.mean())


def test_fullsize_5000_steps_cd():
    """The workload's 5000 steps at 512x256x256, mean Cd over steps 3001-5000. The wake is still developing
    (one flow-through of the domain is 6400 steps): two FP32 runs from different noise seeds agree to 0.1 %,
    while FP16 storage shifts the developing wake by up to about 1 % (measured 0.30-1.04 %); the bound here is
    2 %."""
    c32a, c32b = _cd_window(A.AERO_STORE_FP32, 7), _cd_window(A.AERO_STORE_FP32, 8)
    c16 = _cd_window(A.AERO_STORE_FP16, 7)
    d32 = abs(c32a - c32b) / c32b
    d16 = abs(c16 - c32a) / c32a
    print(f"512x256x256, steps 3001-5000: Cd fp32 seed 7 {c32a:.5f}, seed 8 {c32b:.5f} ({d32 * 100:.3f} %); "
          f"fp16 seed 7 {c16:.5f} ({d16 * 100:.3f} % from fp32)")
    assert d32 < 1e-3
    assert d16 < 0.02
