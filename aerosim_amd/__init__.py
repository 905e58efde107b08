"""MI355X-native D3Q19 wind-tunnel solver (HIP kernels for gfx950 behind the C ABI of include/aerolbm.h)."""
from .lbm import (  # noqa: F401
    AERO_STORE_FP16, AERO_STORE_FP32, AERO_PROFILE_LOGLAW, AERO_PROFILE_UNIFORM, CELL_BOUNDARY, CELL_FLUID,
    CELL_SOLID, AeroError, Batch, Closure, Context, ForceConvergenceTracker, LBMEngine, LBMSolver, SolverParams, Tunnel,
    VoxelGrid, compute_q_criterion, lib, load_stl, mesh_box, mesh_icosphere, mesh_naca, mesh_rocket, mesh_stats,
    normalize_and_center_geometry, parse_stl, voxelize_mesh_to_lattice,
)
