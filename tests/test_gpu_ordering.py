"""Stream ordering and replayed step graphs.

Runs of plain steps replay a captured pair of steps, kept per A-A state and captured again when the arguments
change; batched tunnels step on the batch's own stream, ordered against each tunnel's stream; host-side copies
into a context wait for its pending steps. Each check compares against a run that cannot be affected (one step
per call, explicit synchronisation, single tunnels), bit for bit.
"""
import numpy as np
import pytest

import aerosim_amd as A
from tests.helpers import sphere_case

pytestmark = pytest.mark.gpu


def _sphere(storage=A.AERO_STORE_FP32, u=0.05, seed=3):
    tris, H = sphere_case(64, 32, 32)
    ctx = A.Context(64, 32, 32, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(64.0, 32.0, H))
    ctx.set_params(A.SolverParams(u_inlet=u, nu=0.004, smagorinsky_cs=0.17))
    ctx.reset(seed, 1e-4)
    return ctx


@pytest.mark.parametrize("storage", [A.AERO_STORE_FP32, A.AERO_STORE_FP16])
def test_split_runs_equal_one_run(storage):
    """step(a) then step(b) equals step(a + b), whatever the parities of a and b (graph replays from both A-A
    states, trailing odd steps, a parameter change in between recaptures)."""
    one, parts = _sphere(storage), _sphere(storage)
    one.step(29)
    for n in (1, 5, 2, 7, 6, 8):
        parts.step(n)
    assert one.step_index() == parts.step_index() == 29
    # new parameters: the cached graphs hold the old ones and must not be replayed
    p = A.SolverParams(u_inlet=0.04, nu=0.003, smagorinsky_cs=0.12)
    one.set_params(p)
    parts.set_params(p)
    one.step(12, readout=True)
    parts.step(3)
    parts.step(4)
    parts.step(5, readout=True)
    assert parts.step_index() == 41
    assert np.array_equal(one.read_populations(), parts.read_populations())
    one.close()
    parts.close()


def test_single_steps_equal_graph_replays():
    a, b = _sphere(), _sphere()
    for _ in range(20):
        a.step(1)
    b.step(20)
    a.step(1, readout=True)
    b.step(1, readout=True)
    assert np.array_equal(a.read_populations(), b.read_populations())
    for u, v in zip(a.read_macro(), b.read_macro()):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("pop", [False, True])
def test_frame_loop_graph_equals_single_steps(pop):
    """Frames of 4 substeps, the last with a readout: a captured plain pair then a captured plain + readout pair,
    replayed frame after frame with reads in between, equal single eager steps bit for bit on every frame."""
    a, b = _sphere(), _sphere()
    kw = {"readout": True} if pop else {"macros": True}
    for frame in range(5):
        for _ in range(3):
            a.step(1)
        a.step(1, **kw)
        b.step(4, **kw)
        assert a.step_index() == b.step_index() == 4 * (frame + 1)
        for u, v in zip(a.read_macro(), b.read_macro()):
            assert np.array_equal(u, v), frame
        if pop:
            assert np.array_equal(a.read_populations(), b.read_populations()), frame
        assert a.read_forces()["fx"] == b.read_forces()["fx"], frame
    a.step(1, readout=True)
    b.step(1, readout=True)
    assert np.array_equal(a.read_populations(), b.read_populations())
    a.close()
    b.close()


def test_write_populations_waits_for_pending_steps():
    """A restart written straight after queued steps (no sync) equals the same restart after a sync."""
    src = _sphere()
    src.step(10, readout=True)
    f = src.read_populations()
    a, b = _sphere(), _sphere()
    a.step(200)  # queued, not waited for
    a.write_populations(f)
    a.set_step_index(10)
    b.step(200)
    b.sync()
    b.write_populations(f)
    b.set_step_index(10)
    for c in (a, b):
        c.step(9, readout=True)
    assert a.step_index() == b.step_index() == 19
    assert np.array_equal(a.read_populations(), b.read_populations())


def _naca(nx, ny, nz, alpha, t, chord, Re, u=0.06):
    tris = A.mesh_naca(t, chord, (-0.3 * nx, 0.0, 0.0), alpha, -nz, nz, 48)
    ctx = A.Context(nx, ny, nz)
    ctx.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
    ctx.set_params(A.SolverParams(u_inlet=u, nu=u * chord / Re, smagorinsky_cs=0.17))
    ctx.reset(5, 1e-4)
    return ctx


CASES = [(48, 24, 16, -4.0, 0.12, 12.0, 5e3), (64, 32, 24, 8.0, 0.18, 16.0, 2e4), (40, 24, 24, 16.0, 0.08, 10.0, 5e4)]


def test_batch_then_last_tunnel_first():
    """Per-tunnel reads straight after a batched step, starting with the last tunnel (whose own stream is not the
    batch's), see the batch's finished state; repeated batched calls reuse their tables and graphs."""
    singles = [_naca(*c) for c in CASES]
    batched = [_naca(*c) for c in CASES]
    for c in singles:
        c.step(23)
    b = A.Batch(batched)
    b.set_shape(256, 1, 0)
    for n in (7, 6, 10):
        b.step(n)
    for cs, cb in reversed(list(zip(singles, batched))):
        Fs, Fb = cs.read_forces(), cb.read_forces()
        assert (Fs["fx"], Fs["fy"], Fs["mz"]) == (Fb["fx"], Fb["fy"], Fb["mz"])
    for cs, cb in reversed(list(zip(singles, batched))):
        cs.step(1, readout=True)
        cb.step(1, readout=True)
        assert np.array_equal(cs.read_populations(), cb.read_populations())
    # single-tunnel steps queued before a batched step are waited for by it
    for c in singles:
        c.step(5)
    for c in batched:
        c.step(3)
    b.step(2)
    for cs, cb in zip(singles, batched):
        cs.step(1, readout=True)
        cb.step(1, readout=True)
        assert np.array_equal(cs.read_populations(), cb.read_populations())
    b.close()
