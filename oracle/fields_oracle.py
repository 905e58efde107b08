"""CPU oracle of the derived-field readouts (TEST INFRASTRUCTURE ONLY), in numpy with the reference's
expression order:

- slice_rgba_cpu: extractSliceRGBA with sampleCFDColormap (src/solver/lbmCpuReference.ts:23-46, 500-596), in
  float64 as its JavaScript runs, Math.round halves toward +infinity, NaN stored as 0;
- slice_rgba_wgsl: the post-process shader (src/solver/shaders/postProcess.wgsl:26-146) in float32, with
  WGSL's mix(a, b, t) = a * (1 - t) + b * t, clamp as min(max()) ignoring NaN, truncating u32();
- q_criterion: computeQCriterion (src/render/vortexIso.ts:22-119) in float64, stored to float32;
- probe: the point probe's readout (src/render/pointProbe.ts:284-355), sampleVelocityTrilinear
  (src/render/streamlines.ts:20-81) and computeProbePhysics (pointProbe.ts:23-92), one point at a time.

Pinned by tests/test_fields_oracle.py: the colour map's five stops, ECMAScript rounding, Cp at the ends of
the colour range, Q and the vorticity of linear velocity fields (exact for central and one-sided
differences), the reference's phase5 tests (solid-body rotation, simple shear, probe physics, trilinear
sampling) and trilinear reproduction of linear fields.
"""
import math

import numpy as np

SOLID = 1


def js_round(x):
    x = np.asarray(x, dtype=np.float64)
    r = np.floor(x)
    return np.where(x - r >= 0.5, r + 1.0, r)


def to_u8(x):
    """Uint8Array store: NaN and infinities to 0, otherwise truncation modulo 256."""
    x = np.asarray(x, dtype=np.float64)
    ok = np.isfinite(x)
    t = np.trunc(np.where(ok, x, 0.0)).astype(np.int64)
    return (t & 255).astype(np.uint8)


def colormap64(t):
    """sampleCFDColormap (lbmCpuReference.ts:23-46) on an array of t: (..., 3) float64."""
    t = np.maximum(0.0, np.minimum(1.0, np.asarray(t, dtype=np.float64)))
    out = np.empty(t.shape + (3,))
    stops = [((0.08, 0.22, 0.75), (0.06, 0.72, 0.88)), ((0.06, 0.72, 0.88), (0.12, 0.82, 0.38)),
             ((0.12, 0.82, 0.38), (0.98, 0.62, 0.08)), ((0.98, 0.62, 0.08), (0.94, 0.16, 0.18))]
    seg = np.where(t < 0.25, 0, np.where(t < 0.5, 1, np.where(t < 0.75, 2, 3)))
    base = np.array([0.0, 0.25, 0.5, 0.75])
    for s, (a, b) in enumerate(stops):
        m = seg == s
        u = (t[m] - base[s]) * 4.0 if s else t[m] * 4.0
        for k in range(3):
            out[m, k] = a[k] + (b[k] - a[k]) * u
    return out


def _plane(nx, ny, nz, axis, fixed):
    if axis == 0:
        w, h = nx, ny
        v, u = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        ix, iy, iz = u, v, np.full_like(u, fixed)
    elif axis == 1:
        w, h = nx, nz
        v, u = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        ix, iy, iz = u, np.full_like(u, fixed), v
    else:
        w, h = nz, ny
        v, u = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        ix, iy, iz = np.full_like(u, fixed), v, u
    return w, h, ix, iy, iz


def slice_index(fraction, n):
    return int(min(n - 1, max(0, math.floor(fraction * n))))


def _neighbours(nx, ny, nz, ix, iy, iz):
    idx = lambda x, y, z: x + nx * (y + ny * z)
    return (idx(np.minimum(nx - 1, ix + 1), iy, iz), idx(np.maximum(0, ix - 1), iy, iz),
            idx(ix, np.minimum(ny - 1, iy + 1), iz), idx(ix, np.maximum(0, iy - 1), iz),
            idx(ix, iy, np.minimum(nz - 1, iz + 1)), idx(ix, iy, np.maximum(0, iz - 1)))


def slice_rgba_cpu(nx, ny, nz, ux, uy, uz, rho, flags, axis, fraction, field, u_in):
    """extractSliceRGBA: (height, width, 4) uint8. axis 0/1/2 = 'z'/'y'/'x', field 0/1/2 = velocity/vorticity/cp."""
    n_axis = (nz, ny, nx)[axis]
    w, h, ix, iy, iz = _plane(nx, ny, nz, axis, slice_index(fraction, n_axis))
    idx = ix + nx * (iy + ny * iz)
    ux, uy, uz, rho = (np.asarray(a, dtype=np.float32).astype(np.float64) for a in (ux, uy, uz, rho))
    u_ref = max(1e-4, u_in)
    alpha = np.full(idx.shape, 215.0)
    if field == 0:
        speed = np.sqrt(ux[idx] * ux[idx] + uy[idx] * uy[idx] + uz[idx] * uz[idx])
        scalar = np.minimum(1.0, np.maximum(0.0, speed / (u_ref * 1.55)))
    elif field == 1:
        XP, XM, YP, YM, ZP, ZM = _neighbours(nx, ny, nz, ix, iy, iz)
        wx = 0.5 * (uz[YP] - uz[YM]) - 0.5 * (uy[ZP] - uy[ZM])
        wy = 0.5 * (ux[ZP] - ux[ZM]) - 0.5 * (uz[XP] - uz[XM])
        wz = 0.5 * (uy[XP] - uy[XM]) - 0.5 * (ux[YP] - ux[YM])
        vort = np.sqrt(wx * wx + wy * wy + wz * wz)
        vn = np.minimum(1.0, vort / (u_ref * 0.42))
        scalar = 0.35 + 0.65 * vn
        alpha = js_round(70 + 175 * vn)
    else:
        cp = ((rho[idx] - 1.0) / 3.0) / (0.5 * u_ref * u_ref)
        scalar = np.minimum(1.0, np.maximum(0.0, (cp + 1.1) / 2.1))
    rgb = colormap64(scalar)
    out = np.zeros((h, w, 4), dtype=np.uint8)
    for k in range(3):
        out[..., k] = to_u8(js_round(rgb[..., k] * 255))
    out[..., 3] = to_u8(alpha)
    if flags is not None:
        out[np.asarray(flags)[idx] == SOLID] = 0
    return out


def _f(x):
    return np.float32(x)


def slice_rgba_wgsl(nx, ny, nz, ux, uy, uz, rho, flags, axis, fraction, field, u_in):
    """postProcess.wgsl main(): (height, width, 4) uint8, float32 arithmetic."""
    n_axis = (nz, ny, nx)[axis]
    w, h, ix, iy, iz = _plane(nx, ny, nz, axis, slice_index(fraction, n_axis))
    idx = ix + nx * (iy + ny * iz)
    ux, uy, uz, rho = (np.asarray(a, dtype=np.float32) for a in (ux, uy, uz, rho))
    u_ref = np.fmax(_f(1e-4), _f(u_in))
    clamp = lambda x, lo, hi: np.fmin(np.fmax(x, _f(lo)), _f(hi))
    alpha = np.full(idx.shape, _f(0.82), dtype=np.float32)
    if field == 0:
        x, y, z = ux[idx], uy[idx], uz[idx]
        ratio = np.sqrt(x * x + y * y + z * z) / u_ref
        scalar = clamp(ratio / _f(1.6), 0.0, 1.0)
    elif field == 1:
        XP, XM, YP, YM, ZP, ZM = _neighbours(nx, ny, nz, ix, iy, iz)
        half = _f(0.5)
        dUzDy = half * (uz[YP] - uz[YM])
        dUyDz = half * (uy[ZP] - uy[ZM])
        dUxDz = half * (ux[ZP] - ux[ZM])
        dUzDx = half * (uz[XP] - uz[XM])
        dUyDx = half * (uy[XP] - uy[XM])
        dUxDy = half * (ux[YP] - ux[YM])
        cx, cy, cz = dUzDy - dUyDz, dUxDz - dUzDx, dUyDx - dUxDy
        vn = np.sqrt(cx * cx + cy * cy + cz * cz) / (u_ref * _f(0.45))
        scalar = clamp(_f(0.35) + _f(0.65) * vn, 0.0, 1.0)
        alpha = clamp(_f(0.25) + _f(0.70) * vn, 0.25, 0.92)
    else:
        cp = ((rho[idx] - _f(1.0)) / _f(3.0)) / (_f(0.5) * u_ref * u_ref)
        scalar = clamp((cp + _f(1.1)) / _f(2.1), 0.0, 1.0)
    t = clamp(scalar.astype(np.float32), 0.0, 1.0)
    stops = np.array([[0.08, 0.22, 0.75], [0.06, 0.72, 0.88], [0.12, 0.82, 0.38], [0.98, 0.62, 0.08],
                      [0.94, 0.16, 0.18]], dtype=np.float32)
    seg = np.where(t < _f(0.25), 0, np.where(t < _f(0.5), 1, np.where(t < _f(0.75), 2, 3)))
    base = np.array([0.0, 0.25, 0.5, 0.75], dtype=np.float32)
    u = np.where(seg == 0, t * _f(4.0), (t - base[seg]) * _f(4.0)).astype(np.float32)
    a, b = stops[seg], stops[seg + 1]
    one = _f(1.0)
    rgb = a * (one - u)[..., None] + b * u[..., None]
    pack = lambda x: np.fmin(np.fmax(x * _f(255.0), _f(0.0)), _f(255.0)).astype(np.uint32).astype(np.uint8)
    out = np.zeros((h, w, 4), dtype=np.uint8)
    for k in range(3):
        out[..., k] = pack(rgb[..., k])
    out[..., 3] = pack(alpha)
    if flags is not None:
        out[np.asarray(flags)[idx] == SOLID] = 0
    return out


def q_criterion(nx, ny, nz, dx, ux, uy, uz, cell_flags=None):
    """computeQCriterion: dict of q_values, omega_mag, omega_vec (N x 3), max_q, max_omega."""
    N = nx * ny * nz
    ux, uy, uz = (np.asarray(a, dtype=np.float32).reshape(-1).astype(np.float64) for a in (ux, uy, uz))
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ix, iy, iz = ix.reshape(-1), iy.reshape(-1), iz.reshape(-1)
    inv2dx = 1.0 / (2.0 * dx)

    def scale(prev, nxt):
        d = (nxt - prev).astype(np.float64)
        return np.where(nxt > prev, 1.0 / np.where(nxt > prev, d * dx, 1.0), inv2dx)

    xp, xn = np.maximum(0, ix - 1), np.minimum(nx - 1, ix + 1)
    yp, yn = np.maximum(0, iy - 1), np.minimum(ny - 1, iy + 1)
    zp, zn = np.maximum(0, iz - 1), np.minimum(nz - 1, iz + 1)
    sx, sy, sz = scale(xp, xn), scale(yp, yn), scale(zp, zn)
    g = lambda x, y, z: x + nx * (y + ny * z)
    iXN, iXP, iYN, iYP, iZN, iZP = g(xn, iy, iz), g(xp, iy, iz), g(ix, yn, iz), g(ix, yp, iz), g(ix, iy, zn), g(ix, iy, zp)
    dux_dx = (ux[iXN] - ux[iXP]) * sx
    dux_dy = (ux[iYN] - ux[iYP]) * sy
    dux_dz = (ux[iZN] - ux[iZP]) * sz
    duy_dx = (uy[iXN] - uy[iXP]) * sx
    duy_dy = (uy[iYN] - uy[iYP]) * sy
    duy_dz = (uy[iZN] - uy[iZP]) * sz
    duz_dx = (uz[iXN] - uz[iXP]) * sx
    duz_dy = (uz[iYN] - uz[iYP]) * sy
    duz_dz = (uz[iZN] - uz[iZP]) * sz
    S11, S22, S33 = dux_dx, duy_dy, duz_dz
    S12 = 0.5 * (dux_dy + duy_dx)
    S13 = 0.5 * (dux_dz + duz_dx)
    S23 = 0.5 * (duy_dz + duz_dy)
    normS = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)
    O12 = 0.5 * (dux_dy - duy_dx)
    O13 = 0.5 * (dux_dz - duz_dx)
    O23 = 0.5 * (duy_dz - duz_dy)
    normO = 2.0 * (O12 * O12 + O13 * O13 + O23 * O23)
    qv = 0.5 * (normO - normS)
    ox, oy, oz = duz_dy - duy_dz, dux_dz - duz_dx, duy_dx - dux_dy
    mag = np.sqrt(ox * ox + oy * oy + oz * oz)
    live = np.ones(N, dtype=bool) if cell_flags is None else np.asarray(cell_flags).reshape(-1) != SOLID
    q = np.where(live, qv, 0.0).astype(np.float32)
    om = np.where(live, mag, 0.0).astype(np.float32)
    ov = np.where(live[:, None], np.stack([ox, oy, oz], axis=1), 0.0).astype(np.float32)
    qs, ms = qv[live], mag[live]
    qs, ms = qs[~np.isnan(qs)], ms[~np.isnan(ms)]
    max_q = max(1e-6, float(qs.max())) if qs.size else 1e-6
    max_o = max(1e-6, float(ms.max())) if ms.size else 1e-6
    return {"q_values": q, "omega_mag": om, "omega_vec": ov, "max_q": max_q, "max_omega": max_o}


def _or(x, d):
    x = float(x)
    return d if (x == 0.0 or x != x) else x


def sample_velocity_trilinear(fx, fy, fz, nx, ny, nz, ux, uy, uz):
    """sampleVelocityTrilinear: (ok, (x, y, z))."""
    if fx < 0 or fx >= nx - 1 or fy < 0 or fy >= ny - 1 or fz < 0 or fz >= nz - 1:
        return False, (0.0, 0.0, 0.0)
    ix, iy, iz = math.floor(fx), math.floor(fy), math.floor(fz)
    tx, ty, tz = fx - ix, fy - iy, fz - iz
    i = lambda x, y, z: x + nx * (y + ny * z)
    out = []
    for a in (ux, uy, uz):
        A = lambda x, y, z: float(a[i(x, y, z)])
        c00 = A(ix, iy, iz) * (1 - tx) + A(ix + 1, iy, iz) * tx
        c10 = A(ix, iy + 1, iz) * (1 - tx) + A(ix + 1, iy + 1, iz) * tx
        c01 = A(ix, iy, iz + 1) * (1 - tx) + A(ix + 1, iy, iz + 1) * tx
        c11 = A(ix, iy + 1, iz + 1) * (1 - tx) + A(ix + 1, iy + 1, iz + 1) * tx
        c0 = c00 * (1 - ty) + c10 * ty
        c1 = c01 * (1 - ty) + c11 * ty
        out.append(c0 * (1 - tz) + c1 * tz)
    return True, tuple(out)


def probe_physics(rho, ux, uy, uz, wind, u_in, density, solid):
    """computeProbePhysics: [speed, mph, ux, uy, uz (m/s), gauge Pa, Cp, rho, q_dyn, solid]."""
    q_dyn = 0.5 * density * wind * wind
    q_lat = 0.5 * 1.0 * u_in * u_in
    cp = ((1.0 / 3.0) * (rho - 1.0)) / max(1e-6, q_lat)
    gauge = cp * q_dyn
    if solid:
        return [0.0, 0.0, 0.0, 0.0, 0.0, gauge, cp, rho, q_dyn, 1.0]
    s = wind / max(1e-4, u_in)
    vx, vy, vz = ux * s, uy * s, uz * s
    speed = math.sqrt(vx * vx + vy * vy + vz * vz)
    return [speed, speed * 2.23693629, vx, vy, vz, gauge, cp, rho, q_dyn, 0.0]


def probe(nx, ny, nz, ux, uy, uz, rho, flags, points, origin, dx, wind, u_in, density, on_surface=None):
    """The probe tool's readout at each world point: (n x 13) as aero_probe."""
    out = []
    for n, (px, py, pz) in enumerate(np.asarray(points, dtype=np.float64).reshape(-1, 3)):
        fx = (px - origin[0]) / dx - 0.5
        fy = (py - origin[1]) / dx - 0.5
        fz = (pz - origin[2]) / dx - 0.5
        ix = int(max(0.0, min(nx - 1, float(js_round(fx)))))
        iy = int(max(0.0, min(ny - 1, float(js_round(fy)))))
        iz = int(max(0.0, min(nz - 1, float(js_round(fz)))))
        cell = ix + nx * (iy + ny * iz)
        cell_solid = flags is not None and flags[cell] == SOLID
        solid = bool(on_surface is not None and on_surface[n]) or cell_solid
        ok, v = sample_velocity_trilinear(max(0.0, min(nx - 1.001, fx)), max(0.0, min(ny - 1.001, fy)),
                                          max(0.0, min(nz - 1.001, fz)), nx, ny, nz, ux, uy, uz)
        if (not ok or (v[0] == 0 and v[1] == 0 and v[2] == 0)) and not solid:
            v = (_or(ux[cell], 0.0), _or(uy[cell], 0.0), _or(uz[cell], 0.0))
        r = _or(rho[cell], 1.0)
        if solid and cell_solid:
            for ox, oy, oz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                x, y, z = ix + ox, iy + oy, iz + oz
                if 0 <= x < nx and 0 <= y < ny and 0 <= z < nz:
                    j = x + nx * (y + ny * z)
                    if flags[j] == 0:
                        r = _or(rho[j], r)
                        break
        out.append([ix, iy, iz] + probe_physics(r, v[0], v[1], v[2], wind, u_in, density, solid))
    return np.array(out, dtype=np.float64).reshape(-1, 13)
