"""Pins of the readout oracle (oracle/fields_oracle.py): the colour map's stops, ECMAScript rounding, the
ends of the Cp range, Q and vorticity of linear fields, the reference's phase5 tests (src/validation/
phase5.test.ts:13-132, 181-290) and trilinear sampling. CPU only."""
import math

import numpy as np

from oracle import fields_oracle as F


def test_js_round_and_uint8_store():
    assert F.js_round(2.5) == 3 and F.js_round(-2.5) == -2 and F.js_round(0.49999999999999994) == 0
    assert F.js_round(-0.5) == 0 and F.js_round(1.4999) == 1
    assert F.to_u8([np.nan, np.inf, -1.0, 256.0, 255.9]).tolist() == [0, 0, 255, 0, 255]


def test_colormap_stops():
    """Blue (0) -> cyan (0.25) -> green (0.5) -> orange (0.75) -> red (1): the table's five stops, as bytes."""
    rgb = F.colormap64(np.array([0.0, 0.25, 0.5, 0.75, 1.0, -3.0, 7.0]))
    b = F.to_u8(F.js_round(rgb * 255)).tolist()
    assert b[0] == [20, 56, 191]      # 0.08, 0.22, 0.75
    assert b[1] == [15, 184, 224]     # 0.06, 0.72, 0.88
    assert b[2] == [31, 209, 97]      # 0.12, 0.82, 0.38
    assert b[3] == [250, 158, 20]     # 0.98, 0.62, 0.08
    assert b[4] == [240, 41, 46]      # 0.94, 0.16, 0.18
    assert b[5] == b[0] and b[6] == b[4]  # clamped


def _fields(nx, ny, nz, f):
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return [np.asarray(a, dtype=np.float32).reshape(-1) for a in f(ix.reshape(-1), iy.reshape(-1), iz.reshape(-1))]


def test_slice_cp_ends_and_solid_cells():
    nx, ny, nz, u = 6, 5, 4, 0.08
    N = nx * ny * nz
    z = np.zeros(N, np.float32)
    q = 0.5 * u * u
    for cp, want in ((1.0, [240, 41, 46]), (-1.1, [20, 56, 191]), (5.0, [240, 41, 46]), (-9.0, [20, 56, 191])):
        rho = np.full(N, 1.0 + 3.0 * q * cp, dtype=np.float64).astype(np.float32)
        img = F.slice_rgba_cpu(nx, ny, nz, z, z, z, rho, None, 0, 0.5, 2, u)
        assert img.shape == (ny, nx, 4)
        assert img[2, 3, :3].tolist() == want and img[2, 3, 3] == 215
    flags = np.zeros(N, np.uint8)
    flags[2 + nx * (1 + ny * 2)] = 1  # cell (2, 1, 2) is solid: transparent black in the z = 2 slice
    img = F.slice_rgba_cpu(nx, ny, nz, z, z, z, np.ones(N, np.float32), flags, 0, 0.5, 0, u)
    assert img[1, 2].tolist() == [0, 0, 0, 0] and img[1, 3, 3] == 215


def test_slice_vorticity_of_uniform_flow_and_plane_shapes():
    nx, ny, nz, u = 7, 5, 3, 0.05
    ux, uy, uz = _fields(nx, ny, nz, lambda x, y, z: (np.full(x.shape, u), 0 * x, 0 * x))
    rho = np.ones(nx * ny * nz, np.float32)
    img = F.slice_rgba_cpu(nx, ny, nz, ux, uy, uz, rho, None, 0, 0.5, 1, u)
    assert (img[..., 3] == 70).all()  # zero curl: alpha 70, scalar 0.35
    assert F.slice_rgba_cpu(nx, ny, nz, ux, uy, uz, rho, None, 1, 0.5, 0, u).shape == (nz, nx, 4)
    assert F.slice_rgba_cpu(nx, ny, nz, ux, uy, uz, rho, None, 2, 0.5, 0, u).shape == (ny, nz, 4)
    w = F.slice_rgba_wgsl(nx, ny, nz, ux, uy, uz, rho, None, 0, 0.5, 1, u)
    assert (w[..., 3] == 63).all()    # 0.25 * 255 truncated


def test_slice_index_rule():
    assert F.slice_index(0.5, 64) == 32 and F.slice_index(1.0, 64) == 63 and F.slice_index(-0.2, 64) == 0
    assert F.slice_index(0.999, 10) == 9 and F.slice_index(0.0999, 10) == 0


def test_q_of_linear_fields_is_exact():
    """For u = A x every difference (central inside, one-sided at faces) is exact, so Q and the vorticity are
    the same in every cell: Q = (|W|^2 - |S|^2) / 2 with S, W the parts of A; omega = curl."""
    rng = np.random.default_rng(3)
    nx, ny, nz, dx = 6, 5, 4, 0.25
    A = np.round(rng.standard_normal((3, 3)) * 8) / 8  # dyadic entries keep every step exact
    ux, uy, uz = _fields(nx, ny, nz, lambda x, y, z: tuple(A[i, 0] * x * dx + A[i, 1] * y * dx + A[i, 2] * z * dx
                                                          for i in range(3)))
    r = F.q_criterion(nx, ny, nz, dx, ux, uy, uz)
    S, W = 0.5 * (A + A.T), 0.5 * (A - A.T)
    q = 0.5 * ((W * W).sum() - (S * S).sum())
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.allclose(r["q_values"], q, rtol=0, atol=1e-5)
    assert np.allclose(r["omega_vec"], curl[None, :], atol=1e-5)
    assert abs(r["max_q"] - max(1e-6, q)) < 1e-5 and abs(r["max_omega"] - np.linalg.norm(curl)) < 1e-5


def test_reference_phase5_rotation_and_shear():
    nx = ny = nz = 16
    dx, om0 = 0.1, 5.0
    c = (nx - 1) / 2
    ux, uy, uz = _fields(nx, ny, nz, lambda x, y, z: (-om0 * (y - c) * dx, om0 * (x - c) * dx, 0 * x))
    r = F.q_criterion(nx, ny, nz, dx, ux, uy, uz)
    mid = nx // 2 + nx * (ny // 2 + ny * (nz // 2))
    assert 20.0 < r["q_values"][mid] <= 25.5
    assert round(float(r["omega_mag"][mid]), 0) == 10
    ux, uy, uz = _fields(nx, ny, nz, lambda x, y, z: (4.0 * (y * dx), 0 * x, 0 * x))
    assert abs(F.q_criterion(nx, ny, nz, dx, ux, uy, uz)["q_values"][mid]) < 1e-4


def test_solid_cells_are_skipped():
    nx, ny, nz = 5, 4, 3
    ux, uy, uz = _fields(nx, ny, nz, lambda x, y, z: (np.sin(x + 2.0 * y), np.cos(y * z + 0.0), 0.3 * x * z))
    flags = np.zeros(nx * ny * nz, np.uint8)
    flags[7] = 1
    r = F.q_criterion(nx, ny, nz, 1.0, ux, uy, uz, flags)
    assert r["q_values"][7] == 0 and r["omega_mag"][7] == 0 and (r["omega_vec"][7] == 0).all()
    full = F.q_criterion(nx, ny, nz, 1.0, ux, uy, uz)
    keep = np.arange(nx * ny * nz) != 7
    assert np.array_equal(r["q_values"][keep], full["q_values"][keep])


def test_reference_phase5_probe_physics():
    u_in, wind = 0.085, 30.0
    q_lat = 0.5 * u_in * u_in
    stag = 1.0 + q_lat / (1.0 / 3.0)
    p = F.probe_physics(stag, 0.0, 0.0, 0.0, wind, u_in, 1.204, False)
    assert p[0] == 0 and p[1] == 0 and abs(p[6] - 1.0) < 5e-4
    assert abs(p[5] - 0.5 * 1.204 * wind * wind) < 0.05
    s = F.probe_physics(1.0, 0.08, 0.0, 0.0, 25.0, u_in, 1.204, True)
    assert s[9] == 1.0 and s[0] == 0 and s[5] == 0


def test_trilinear_sampling():
    nx = ny = nz = 4
    N = nx * ny * nz
    ok, v = F.sample_velocity_trilinear(1.5, 1.5, 1.5, nx, ny, nz, np.full(N, 10.0, np.float32),
                                        np.full(N, 2.0, np.float32), np.full(N, -1.0, np.float32))
    assert ok and abs(v[0] - 10) < 1e-4 and abs(v[1] - 2) < 1e-4 and abs(v[2] + 1) < 1e-4
    ok, v = F.sample_velocity_trilinear(-0.5, 2.0, 2.0, nx, ny, nz, *(np.ones(N, np.float32),) * 3)
    assert not ok and v == (0.0, 0.0, 0.0)
    # a linear field is reproduced exactly (dyadic coefficients and sample points)
    ux, uy, uz = _fields(nx, ny, nz, lambda x, y, z: (0.5 * x - 0.25 * y + 2.0 * z, 1.0 + x, y * 0.125 - z))
    for p in ((0.25, 1.5, 2.75), (2.0, 0.0, 1.125), (2.875, 2.5, 0.5)):
        ok, v = F.sample_velocity_trilinear(*p, nx, ny, nz, ux, uy, uz)
        want = (0.5 * p[0] - 0.25 * p[1] + 2.0 * p[2], 1.0 + p[0], p[1] * 0.125 - p[2])
        assert ok and v == want


def test_probe_nearest_fluid_density_on_solid_hits():
    nx, ny, nz = 6, 6, 6
    N = nx * ny * nz
    flags = np.zeros(N, np.uint8)
    flags[3 + nx * (3 + ny * 3)] = 1
    flags[4 + nx * (3 + ny * 3)] = 2          # +x neighbour is a boundary cell: skipped
    rho = np.ones(N, np.float32)
    rho[2 + nx * (3 + ny * 3)] = 1.02         # -x neighbour: the first fluid one in the search order
    z = np.zeros(N, np.float32)
    out = F.probe(nx, ny, nz, z, z, z, rho, flags, [[3.5, 3.5, 3.5]], (0, 0, 0), 1.0, 25.0, 0.085, 1.204)
    assert out[0, :3].tolist() == [3, 3, 3] and out[0, 12] == 1.0
    assert out[0, 10] == float(np.float32(1.02))
    assert math.isclose(out[0, 9], (1 / 3) * (float(np.float32(1.02)) - 1) / (0.5 * 0.085 ** 2))


def test_reference_phase5_streamline_advection():
    """phase5.test.ts:205-240: a streamline traced with the trilinear sampler through uniform flow along +x
    (12 x 8 x 8, dx = 0.2, 15 half-cell steps) moves downstream and keeps its y and z."""
    nx, ny, nz, dx = 12, 8, 8, 0.2
    ux = np.full(nx * ny * nz, 0.08, dtype=np.float32)
    uy = np.zeros_like(ux)
    uz = np.zeros_like(ux)
    x, y, z = 0.5 * dx, 4.0 * dx, 4.0 * dx
    ds = dx * 0.5
    for _ in range(15):
        ok, v = F.sample_velocity_trilinear(x / dx - 0.5, y / dx - 0.5, z / dx - 0.5, nx, ny, nz, ux, uy, uz)
        assert ok
        spd = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        x += ds * v[0] / spd
        y += ds * v[1] / spd
        z += ds * v[2] / spd
    assert x > 1.0
    assert abs(y - 4.0 * dx) < 5e-5 and abs(z - 4.0 * dx) < 5e-5
