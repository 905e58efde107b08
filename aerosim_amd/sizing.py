"""Grid sizing and force/coefficient conversions of the reference's host, restated with its JavaScript
arithmetic: calculateSafeGridResolution (src/core/gpuContext.ts:10-58) and forceToCoefficient /
coefficientToForce (src/solver/fluidProperties.ts:54-80).

calculate_safe_grid_resolution keeps the reference's inputs, the largest single storage buffer and the largest
buffer a device allows (bytes), the tunnel and a cell budget, and its output, including the reference's
estimate of 216 B per cell. On an MI355X the populations live in one A-A buffer of 19 values per cell (76 B in
FP32, 38 B in FP16), and the solver accepts grids below 2^31 cells, so `mi355x_buffer_limits` gives limits
from the GPU's memory rather than a WebGPU adapter's.
"""
import math

import numpy as np

from .export import js_to_fixed

BYTES_PER_LBM_CELL = 216  # gpuContext.ts:10: the reference's per-cell budget for its estimate


def _js_round(x):
    """Math.round: halves toward +infinity."""
    r = math.floor(x)
    return r + 1 if x - r >= 0.5 else r


def calculate_safe_grid_resolution(max_storage_buffer_binding_size, max_buffer_size, tunnel, target_max_cells):
    """calculateSafeGridResolution (gpuContext.ts:17-58): the largest grid that keeps one 19-value FP32
    population buffer under both limits (with margin) and under the cell budget, in the tunnel's aspect ratio,
    every size a multiple of 8 (nx >= 64, ny, nz >= 32). tunnel: object or dict with length, width, height.
    Returns dict(nx, ny, nz, totalCells, estimatedVramMB)."""
    get = (lambda k: tunnel[k]) if isinstance(tunnel, dict) else (lambda k: getattr(tunnel, k))
    per_cell = 19 * 4
    limit = min(max_storage_buffer_binding_size, max_buffer_size)
    hardware_max = math.floor(limit / per_cell)
    safe = max(32 * 16 * 16, min(hardware_max * 0.85, target_max_cells))
    lx, ly, lz = max(0.1, get("length")), max(0.1, get("height")), max(0.1, get("width"))
    pitch = float(np.cbrt((lx * ly * lz) / safe))

    def align8(n, lo):
        return max(lo, math.floor(n / 8) * 8)

    ny = align8(_js_round(ly / pitch), 32)
    h = ly / ny
    nx = align8(_js_round(lx / h), 64)
    nz = align8(_js_round(lz / h), 32)
    while nx * ny * nz * per_cell > limit * 0.9:
        ny = max(32, ny - 8)
        nh = ly / ny
        nx = max(64, math.floor((lx / nh) / 8) * 8)
        nz = max(32, math.floor((lz / nh) / 8) * 8)
    total = nx * ny * nz
    return {"nx": nx, "ny": ny, "nz": nz, "totalCells": total,
            "estimatedVramMB": float(js_to_fixed((total * BYTES_PER_LBM_CELL) / (1024 * 1024), 1))}


def mi355x_buffer_limits(hbm_bytes=288 * 1024 ** 3, fraction=0.5, storage_bytes=4):
    """Limits to pass as the first two arguments of calculate_safe_grid_resolution on an MI355X: the
    population buffer may take `fraction` of the HBM, and the grid stays below 2^31 cells."""
    cap = min(int(hbm_bytes * fraction), (2 ** 31 - 1) * 19 * storage_bytes)
    return cap, cap


def force_to_coefficient(force_newtons, density, wind_speed, reference_area_m2):
    """forceToCoefficient (fluidProperties.ts:58-67): C = 2 F / (rho U^2 A), 0 for a vanishing denominator."""
    denom = 0.5 * density * wind_speed * wind_speed * reference_area_m2
    if denom <= 1e-12:
        return 0.0
    return force_newtons / denom


def coefficient_to_force(coefficient, density, wind_speed, reference_area_m2):
    """coefficientToForce (fluidProperties.ts:73-80): F = rho U^2 A C / 2."""
    return 0.5 * density * wind_speed * wind_speed * reference_area_m2 * coefficient
