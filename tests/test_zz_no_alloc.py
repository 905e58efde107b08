"""aero_read_forces_noalloc, the force readout that allocates nothing, against aero_read_forces through the C ABI:
bit for bit on two calls in a row, the second on a context with fewer boundary cells (the shared per-GPU scratch
keeps the first call's partials, which the second must not pick up), and again after further steps."""
import numpy as np
import pytest

import aerosim_amd as A
from tests.helpers import rocket_case, sphere_case

pytestmark = pytest.mark.gpu


def _ctx(tris, dims, H, storage=A.AERO_STORE_FP32):
    ctx = A.Context(*dims, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(dims[0]), float(dims[2]), H))
    ctx.set_params(A.SolverParams(u_inlet=0.06, nu=0.003, smagorinsky_cs=0.17))
    ctx.reset(17, 1e-4)
    return ctx


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_noalloc_forces_equal_allocating_forces(storage):
    tris, H = rocket_case(256, 128, 128, scale=0.5)
    big = _ctx(tris, (256, 128, 128), H, storage)
    tris_s, Hs = sphere_case(64, 32, 32)
    small = _ctx(tris_s, (64, 32, 32), Hs, storage)
    nb_big, nb_small = big.read_voxels().boundary_cell_count, small.read_voxels().boundary_cell_count
    print(f"boundary cells: {nb_big} then {nb_small}")
    assert nb_big > 4 * nb_small > 0
    refs = dict(frontal_area=100.0, char_length=80.0, tunnel_width=128.0, tunnel_height=128.0, dx=1.0, u_ref=0.06)
    for ctx, n in ((big, 30), (small, 30), (big, 7), (small, 9)):
        ctx.step(n)
        a = ctx.read_forces(refs)
        b = ctx.read_forces(refs, noalloc=True)
        print(a["fx"], b["fx"], a["cd"], b["cd"])
        assert a == b
        assert np.isfinite(a["fx"]) and a["cut_face_links"] > 0
    big.close()
    small.close()
