"""Shared workload builders for the tests (meshes in lattice units, voxels from the oracle)."""
import numpy as np

import aerosim_amd as A


def sphere_case(nx=64, ny=32, nz=32, diameter=8.0, x_centre=20.0, level=3):
    """Icosphere of `diameter` cells centred at lattice x = x_centre, centred in y/z (tunnel height = ny)."""
    tris = A.mesh_icosphere(level, 0.5 * diameter, (x_centre - 0.5 * nx, 0.0, 0.0))
    return tris, float(ny)


def rocket_case(nx=128, ny=64, nz=64, scale=0.25, seed=7):
    """Finned rocket of the 512x256x256 workload scaled by `scale` (nose tip at lattice x = 100*scale)."""
    D, L, Ln = 24.0 * scale, 160.0 * scale, 60.0 * scale
    tip = (100.0 * scale - 0.5 * nx, 0.0, 0.0)
    tris = A.mesh_rocket(tip, L, D, Ln, 0.02, seed, 128)
    return tris, float(ny)


def relerr(a, b, floor=1e-12):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))
