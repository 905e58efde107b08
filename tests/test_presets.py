"""Preset meshes (aerosim_amd/presets.py) through the mesh statistics, with the reference's Phase 1 checks
(src/validation/phase1.test.ts:11-49) and closed forms of the three.js constructions. CPU only."""
import math

import numpy as np

import aerosim_amd as A
from aerosim_amd import presets as P


def test_phase1_sphere_frontal_area():
    """phase1.test.ts:11-20: within 0.5 % of pi r^2 for the 1 m sphere."""
    st = A.mesh_stats(P.create_preset_geometry("sphere"))
    assert abs(st["frontal_area"] - math.pi * 0.25) / (math.pi * 0.25) < 0.005
    assert round(st["frontal_area"], 2) == 0.79


def test_phase1_cube_frontal_area_and_rotation():
    """phase1.test.ts:22-37: 1.000 m^2 at 0 deg; sqrt(2) m^2 within 0.5 % after 45 deg about y."""
    cube = P.create_preset_geometry("cube")
    assert abs(A.mesh_stats(cube)["frontal_area"] - 1.0) < 5e-4
    c, s = math.cos(math.pi / 4), math.sin(math.pi / 4)
    rot_y = np.array([c, 0, -s, 0, 0, 1, 0, 0, s, 0, c, 0, 0, 0, 0, 1], dtype=np.float64)  # makeRotationY, column-major
    area = A.mesh_stats(cube, rot_y)["frontal_area"]
    assert abs(area - math.sqrt(2)) / math.sqrt(2) < 0.005


def test_phase1_cube_watertight_stats():
    """phase1.test.ts:39-49."""
    st = A.mesh_stats(P.create_preset_geometry("cube"))
    assert st["watertight"] and st["open_edges"] == 0
    assert abs(st["volume"] - 1.0) < 5e-4 and abs(st["frontal_area"] - 1.0) < 5e-4


def test_three_js_construction_counts_and_closed_forms():
    sphere = P.create_preset_geometry("sphere")
    assert sphere.shape == (96 * 64 * 2 - 2 * 96, 9)          # no triangles at the poles' degenerate rows
    r = np.linalg.norm(sphere.reshape(-1, 3).astype(np.float64), axis=1)
    assert np.allclose(r, 0.5, atol=1e-7)
    cyl = P.create_preset_geometry("cylinder")
    assert cyl.shape == (80 * 4 * 2 + 2 * 80, 9)
    v = cyl.reshape(-1, 3).astype(np.float64)
    assert np.allclose(np.abs(v[:, 1]).max(), 0.6) and A.mesh_stats(cyl)["watertight"]
    assert abs(A.mesh_stats(cyl)["frontal_area"] - 0.72) < 2e-3
    nose = P.create_preset_geometry("rocket_nose")
    v = nose.reshape(-1, 3).astype(np.float64)
    assert np.allclose(v.min(0), [-0.7, -0.25, -0.25], atol=1e-6) and np.allclose(v.max(0), [0.7, 0.25, 0.25], atol=1e-6)
    tip = v[np.argmin(v[:, 0])]
    assert abs(tip[1]) < 1e-6 and abs(tip[2]) < 1e-6           # the tip points to -x, on the axis
    assert abs(A.mesh_stats(nose)["frontal_area"] - math.pi * 0.25 ** 2) / (math.pi * 0.25 ** 2) < 0.01


def test_shed_closed_forms():
    """The gable-roof shed (presets.ts:113-138): 0.9 x 0.95 x 1.2 m, watertight, volume 0.9 x 0.65 x 1.2 + the roof
    prism 0.9 x 0.3 / 2 x 1.2 = 0.864 m^3, frontal area (projection along the wind) 1.2 x 0.95 = 1.14 m^2,
    centred on its bounding box."""
    t = P.create_preset_geometry("shed")
    st = A.mesh_stats(t)
    assert st["watertight"] and st["open_edges"] == 0
    assert abs(st["volume"] - 0.864) < 1e-6 and abs(st["frontal_area"] - 1.14) < 5e-4
    v = t.reshape(-1, 3).astype(np.float64)
    assert np.allclose(v.min(0), [-0.45, -0.475, -0.6], atol=1e-7) and np.allclose(v.max(0), [0.45, 0.475, 0.6], atol=1e-7)
    assert abs(st["characteristic_length"] - 0.9) < 1e-6
