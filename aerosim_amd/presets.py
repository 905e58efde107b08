"""Preset specimen meshes of the reference's viewer (createPresetGeometry, src/geometry/presets.ts:59-112) as
triangle soups (n x 9 float32), built the way three.js r170 builds them: SphereGeometry(0.5, 96, 64),
BoxGeometry(1, 1, 1, 4, 4, 4), CylinderGeometry(0.3, 0.3, 1.2, 80, 4) and the tangent-ogive LatheGeometry with 72
segments turned by rotateZ(-pi/2) and centred. Vertex coordinates are computed in double and stored as float,
as the geometries' Float32 position attributes hold them; a matrix transform is applied in double to the
stored floats and stored back to float (Vector3.applyMatrix4). The `shed` preset is the reference's gable
profile extruded as ExtrudeGeometry does it (depth 1.2, 2 steps, no bevel) and centred; its end caps are
triangulated as a fan rather than by three.js's earcut, which bounds the same solid (the voxels, frontal area
and volume depend only on the solid)."""
import math

import numpy as np

PRESET_IDS = ("sphere", "cube", "cylinder", "rocket_nose", "shed")


def _soup(pos, faces):
    p = np.asarray(pos, dtype=np.float64).astype(np.float32)
    f = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    return p[f].reshape(-1, 9)


def sphere_geometry(radius=0.5, width_segments=96, height_segments=64):
    """SphereGeometry with phi in [0, 2 pi) and theta in [0, pi]."""
    pos, grid = [], []
    for iy in range(height_segments + 1):
        v = iy / height_segments
        row = []
        for ix in range(width_segments + 1):
            u = ix / width_segments
            phi, theta = u * 2 * math.pi, v * math.pi
            pos.append((-radius * math.cos(phi) * math.sin(theta), radius * math.cos(theta),
                        radius * math.sin(phi) * math.sin(theta)))
            row.append(len(pos) - 1)
        grid.append(row)
    faces = []
    for iy in range(height_segments):
        for ix in range(width_segments):
            a, b, c, d = grid[iy][ix + 1], grid[iy][ix], grid[iy + 1][ix], grid[iy + 1][ix + 1]
            if iy != 0:
                faces.append((a, b, d))
            if iy != height_segments - 1:
                faces.append((b, c, d))
    return _soup(pos, faces)


def box_geometry(width=1.0, height=1.0, depth=1.0, ws=4, hs=4, ds=4):
    """BoxGeometry: six segmented planes (+x, -x, +y, -y, +z, -z)."""
    pos, faces = [], []
    axis = {"x": 0, "y": 1, "z": 2}

    def plane(u, v, w, udir, vdir, pw, ph, pd, gx, gy):
        start = len(pos)
        sw, sh = pw / gx, ph / gy
        for iy in range(gy + 1):
            y = iy * sh - ph / 2
            for ix in range(gx + 1):
                x = ix * sw - pw / 2
                vec = [0.0, 0.0, 0.0]
                vec[axis[u]], vec[axis[v]], vec[axis[w]] = x * udir, y * vdir, pd / 2
                pos.append(tuple(vec))
        g1 = gx + 1
        for iy in range(gy):
            for ix in range(gx):
                a, b = start + ix + g1 * iy, start + ix + g1 * (iy + 1)
                c, d = start + (ix + 1) + g1 * (iy + 1), start + (ix + 1) + g1 * iy
                faces.extend([(a, b, d), (b, c, d)])

    plane("z", "y", "x", -1, -1, depth, height, width, ds, hs)
    plane("z", "y", "x", 1, -1, depth, height, -width, ds, hs)
    plane("x", "z", "y", 1, 1, width, depth, height, ws, ds)
    plane("x", "z", "y", 1, -1, width, depth, -height, ws, ds)
    plane("x", "y", "z", 1, -1, width, height, depth, ws, hs)
    plane("x", "y", "z", -1, -1, width, height, -depth, ws, hs)
    return _soup(pos, faces)


def cylinder_geometry(radius_top=0.3, radius_bottom=0.3, height=1.2, radial_segments=80, height_segments=4):
    """CylinderGeometry along y, with both caps."""
    pos, faces, index = [], [], []
    half = height / 2
    for y in range(height_segments + 1):
        v = y / height_segments
        r = v * (radius_bottom - radius_top) + radius_top
        row = []
        for x in range(radial_segments + 1):
            th = (x / radial_segments) * 2 * math.pi
            pos.append((r * math.sin(th), -v * height + half, r * math.cos(th)))
            row.append(len(pos) - 1)
        index.append(row)
    for x in range(radial_segments):
        for y in range(height_segments):
            a, b, c, d = index[y][x], index[y + 1][x], index[y + 1][x + 1], index[y][x + 1]
            if radius_top > 0 or y != 0:
                faces.append((a, b, d))
            if radius_bottom > 0 or y != height_segments - 1:
                faces.append((b, c, d))
    for top in (True, False):
        r, sign = (radius_top, 1) if top else (radius_bottom, -1)
        c0 = len(pos)
        for _ in range(radial_segments):
            pos.append((0.0, half * sign, 0.0))
        e0 = len(pos)
        for x in range(radial_segments + 1):
            th = (x / radial_segments) * 2 * math.pi
            pos.append((r * math.sin(th), half * sign, r * math.cos(th)))
        for x in range(radial_segments):
            c, i = c0 + x, e0 + x
            faces.append((i, i + 1, c) if top else (i + 1, i, c))
    return _soup(pos, faces)


def lathe_geometry(points, segments):
    """LatheGeometry: the (radius, axial) profile revolved about y."""
    pos = []
    n = len(points)
    for i in range(segments + 1):
        phi = i * (1.0 / segments) * 2 * math.pi
        s, c = math.sin(phi), math.cos(phi)
        for px, py in points:
            pos.append((px * s, py, px * c))
    faces = []
    for i in range(segments):
        for j in range(n - 1):
            base = j + i * n
            a, b, c, d = base, base + n, base + n + 1, base + 1
            faces.extend([(a, b, d), (c, d, b)])
    return _soup(pos, faces)


def _apply(tris, e):
    """Vector3.applyMatrix4 of a column-major 4x4 on every stored vertex, stored back to float."""
    v = tris.reshape(-1, 3).astype(np.float64)
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    w = 1.0 / (e[3] * x + e[7] * y + e[11] * z + e[15])
    out = np.stack([(e[0] * x + e[4] * y + e[8] * z + e[12]) * w, (e[1] * x + e[5] * y + e[9] * z + e[13]) * w,
                    (e[2] * x + e[6] * y + e[10] * z + e[14]) * w], axis=1)
    return out.astype(np.float32).reshape(-1, 9)


def rocket_nose_geometry():
    """The tangent ogive nose cone (1.0 m ogive, 0.4 m body tube, base radius 0.25 m) pointing to -x, centred."""
    L, body, R = 1.0, 0.4, 0.25
    rho = (R * R + L * L) / (2 * R)
    pts = []
    for i in range(61):
        x = (i / 60) * L
        y = math.sqrt(max(0.0, rho * rho - (L - x) * (L - x))) + R - rho
        pts.append((max(0.0, y), x))
    pts += [(R, L + body), (0.0, L + body)]
    t = lathe_geometry(pts, 72)
    c, s = math.cos(-math.pi / 2), math.sin(-math.pi / 2)
    t = _apply(t, [c, s, 0, 0, -s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])   # rotateZ(-pi/2)
    v = t.reshape(-1, 3).astype(np.float64)
    ctr = (v.min(0) + v.max(0)) * 0.5                                  # center(): translate by -box centre
    return _apply(t, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -ctr[0], -ctr[1], -ctr[2], 1])


def shed_geometry(hw=0.45, wall_h=0.65, ridge_h=0.95, depth=1.2, steps=2):
    """The gable-roof shed (presets.ts:113-138): the pentagon (-hw, 0), (hw, 0), (hw, wall_h), (0, ridge_h),
    (-hw, wall_h) in the x-y plane extruded along +z over `depth` in `steps` slices, outward-facing, then centred on
    its bounding box."""
    shape = [(-hw, 0.0), (hw, 0.0), (hw, wall_h), (0.0, ridge_h), (-hw, wall_h)]
    n = len(shape)
    zs = [depth * s / steps for s in range(steps + 1)]
    pos = [(x, y, z) for z in zs for (x, y) in shape]
    faces = []
    top = steps * n
    for i in range(1, n - 1):  # caps: a fan over the convex profile, facing -z at z = 0 and +z at z = depth
        faces.append((0, i + 1, i))
        faces.append((top, top + i, top + i + 1))
    for s in range(steps):
        for i in range(n):
            j = (i + 1) % n
            a, b, c, d = s * n + i, s * n + j, (s + 1) * n + j, (s + 1) * n + i
            faces.extend([(a, b, c), (a, c, d)])
    t = _soup(pos, faces)
    v = t.reshape(-1, 3).astype(np.float64)
    ctr = (v.min(0) + v.max(0)) * 0.5
    return _apply(t, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -ctr[0], -ctr[1], -ctr[2], 1])


def create_preset_geometry(preset_id):
    """createPresetGeometry (presets.ts:59-112): triangle soup of a preset specimen (metres, centred)."""
    if preset_id == "sphere":
        return sphere_geometry()
    if preset_id == "cube":
        return box_geometry()
    if preset_id == "cylinder":
        return cylinder_geometry()
    if preset_id == "rocket_nose":
        return rocket_nose_geometry()
    if preset_id == "shed":
        return shed_geometry()
    raise ValueError(f"unknown preset {preset_id!r} (have {', '.join(PRESET_IDS)})")
