"""Error behaviour of the C ABI through the Python host: calls out of order or with bad arguments fail with
the documented AERO_E_* code and a message naming the problem, and leave the context usable."""
import numpy as np
import pytest

import aerosim_amd as A
from aerosim_amd import lbm as L
from tests.helpers import sphere_case

pytestmark = pytest.mark.gpu


def _expect(code, text, fn, *args):
    with pytest.raises(A.AeroError) as e:
        fn(*args)
    assert e.value.code == code, str(e.value)
    assert text in str(e.value), str(e.value)


def test_state_and_argument_errors_leave_the_context_usable():
    tris, H = sphere_case(64, 32, 32)
    ctx = A.Context(64, 32, 32)
    _expect(L.AERO_E_STATE, "no geometry", ctx.reset, 0, 0.0)
    ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    _expect(L.AERO_E_STATE, "aero_reset", ctx.step, 1)
    ctx.set_params(A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17))
    for k in (9, 17, 24, -1):
        _expect(L.AERO_E_ARG, "fp16_shift", ctx.set_params,
                A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17, fp16_shift=k))
    _expect(L.AERO_E_STATE, "closure", ctx.set_params,
            A.SolverParams(u_inlet=0.05, nu=0.004, smagorinsky_cs=0.17, closure=L.AERO_CLOSURE_MLP))
    ctx.reset(1, 1e-4)
    _expect(L.AERO_E_ARG, "negative", ctx.step, -1)
    _expect(L.AERO_E_ARG, "1, 2 or 4", ctx.set_cells_per_thread, 3)
    ctx.step(4)
    _expect(L.AERO_E_STATE, "AERO_STEP_READOUT", ctx.read_populations)
    # the context still steps and reads after the failed calls
    ctx.step(2, readout=True)
    s = ctx.mass_stats()
    assert not s["has_nan"] and abs(s["mean_rho"] - 1.0) < 1e-3
    assert ctx.step_index() == 6
    ctx.close()


def test_cells_per_thread_must_divide_nx():
    tris, H = sphere_case(65, 32, 32)
    ctx = A.Context(65, 32, 32)
    _expect(L.AERO_E_ARG, "dividing nx", ctx.set_cells_per_thread, 2)
    ctx.close()


def test_batch_rejects_mixed_storage():
    tris, H = sphere_case(64, 32, 32)
    a = A.Context(64, 32, 32, storage=A.AERO_STORE_FP32)
    b = A.Context(64, 32, 32, storage=A.AERO_STORE_FP16)
    for c in (a, b):
        c.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    _expect(L.AERO_E_ARG, "another storage", A.Batch, [a, b])
    a.close()
    b.close()


def test_short_voxel_arrays_are_rejected_before_the_library_reads_them():
    tris, H = sphere_case(64, 32, 32)
    ctx = A.Context(64, 32, 32)
    ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    vox = ctx.read_voxels(H)
    keys = ("cell_flags", "boundary_index_map", "boundary_cell_indices", "boundary_q_values")
    for short in ("cell_flags", "boundary_index_map", "boundary_q_values"):
        v = {k: getattr(vox, k) for k in keys}
        v[short] = v[short][:-1]
        _expect(L.AERO_E_ARG, "set_voxels", ctx.set_voxels, type("ShortVoxels", (), v))
    ctx.set_voxels(vox)  # the full arrays still load
    ctx.close()
