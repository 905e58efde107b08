"""The reference's automated wind-tunnel physics benchmark (runSinglePhysicsBenchmark,
src/validation/testRunner.ts:39-150) on the GPU solver, with its empirical targets
(src/validation/analyticalCases.ts).

Protocol: a preset specimen (aerosim_amd.presets) voxelized into a 6.0 x 3.6 x 3.6 m tunnel on 60 x 36 x 36
cells, started from equilibrium at u = 0.001, the inlet ramped to the target speed as sin^2 over 80 steps,
Smagorinsky Cs = 0.12, forces sampled on every step after the first half of the run, and the mean Cd of the
last 40 samples (ForceConvergenceTracker window 40), rounded to 3 decimals, checked against the target band at
the simulated Reynolds number.
"""
import math
from dataclasses import dataclass

from . import lbm as L
from .export import js_to_fixed
from .presets import create_preset_geometry

DEFAULT_SMAGORINSKY_CS = 0.12   # lbmConstants.ts:80

# analyticalCases.ts:21-67: target Cd and band of each preset
EMPIRICAL_BENCHMARK_CASES = {
    "sphere": {"name": "Smooth Calibration Sphere", "target": 0.47, "min": 0.38, "max": 0.56},
    "cylinder": {"name": "Finite Circular Cylinder in Cross-Flow (L/D = 2)", "target": 0.68, "min": 0.58, "max": 0.78},
    "cube": {"name": "Sharp-Edged Bluff Cube", "target": 1.05, "min": 0.88, "max": 1.22},
    "rocket_nose": {"name": "Streamlined Ogive Rocket Nose Cone", "target": 0.19, "min": 0.13, "max": 0.26},
    "shed": {"name": "Gable-Roof Backyard Shed (Ground ABL)", "target": 1.04, "min": 0.86, "max": 1.24},
}


def _fixed(x, d):
    return float(js_to_fixed(x, d))


def analytical_sphere_cd(re):
    """Morrison / Clift-Grace-Weber sphere drag (analyticalCases.ts:73-82)."""
    re = max(0.1, re)
    if re > 3.5e5:
        return 0.20
    return (24.0 / re) * (1.0 + 0.15 * re ** 0.687) + 0.42 / (1.0 + 42500.0 * re ** -1.16)


def analytical_finite_cylinder_cd(re, ld=2):
    """Finite cylinder with free ends (analyticalCases.ts:88-92)."""
    return (0.68 if ld == 2 else 0.75) + 5.5 / math.sqrt(max(10, re))


def analytical_rocket_nose_cd(re):
    """Ogive nose with blunt base (analyticalCases.ts:98-101)."""
    return 0.20 + 26.0 / math.sqrt(max(10, re))


def benchmark_target(preset_id, re):
    """getBenchmarkTargetForSimulatedRe (analyticalCases.ts:106-135): (target, min, max)."""
    if preset_id == "sphere":
        t = _fixed(analytical_sphere_cd(re), 3)
        return t, _fixed(t * 0.78, 3), _fixed(t * 1.25, 3)
    if preset_id == "cylinder":
        t = _fixed(analytical_finite_cylinder_cd(re, 2), 3)
        return t, _fixed(t * 0.80, 3), _fixed(t * 1.20, 3)
    if preset_id == "rocket_nose":
        t = _fixed(analytical_rocket_nose_cd(re), 3)
        return t, _fixed(t * 0.80, 3), _fixed(t * 1.25, 3)
    b = EMPIRICAL_BENCHMARK_CASES[preset_id]
    return b["target"], b["min"], b["max"]


def _js_round(x):
    r = math.floor(x)
    return r + 1 if x - r >= 0.5 else r


@dataclass
class BenchmarkResult:
    preset_id: str
    computed_cd: float
    computed_cl: float
    target_cd: float
    min_cd: float
    max_cd: float
    relative_error_percent: float
    passed: bool
    steps: int
    simulated_re: float
    cells_across_d: float


def run_single_physics_benchmark(preset_id, grid=(60, 36, 36), steps=320, tunnel=None, target_u=0.08,
                                 nu_lattice=0.002, sample_after=0.5, window=40, tris=None, device=0):
    """runSinglePhysicsBenchmark (testRunner.ts:39-150). sample_after: the fraction of the run after which forces
    are sampled (the grid-convergence test uses 0.6); tris overrides the preset mesh."""
    tunnel = tunnel or L.Tunnel(6.0, 3.6, 3.6)
    tris = create_preset_geometry(preset_id) if tris is None else tris
    st = L.mesh_stats(tris)
    vox = L.voxelize_mesh_to_lattice(tris, None, tunnel, grid, device=device)
    lat_d = st["characteristic_length"] / vox.dx
    cells_across_d = _fixed(lat_d, 1)
    re = _js_round((target_u * lat_d) / nu_lattice)
    ctx = L.Context(*grid, device=device)
    try:
        ctx.set_voxels(vox)
        ctx.set_params(L.SolverParams(u_inlet=target_u, nu=nu_lattice, smagorinsky_cs=DEFAULT_SMAGORINSKY_CS,
                                      ground_noslip=preset_id == "shed", ramp_steps=80, u_init=0.001))
        ctx.reset(0, 0.0)
        refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"],
                    tunnel_width=tunnel.width, tunnel_height=tunnel.height, dx=vox.dx, u_ref=target_u)
        tracker = L.ForceConvergenceTracker(window)
        first = math.floor(steps * sample_after)
        ctx.step(first)
        readout = None
        for s in range(first + 1, steps + 1):
            ctx.step(1)
            f = ctx.read_forces(refs)
            readout = tracker.push_sample(s, f["cd"], f["cl"], f["cm"], f["xcp"], 25.0, 1.204, st["frontal_area"])
    finally:
        ctx.close()
    cd = _fixed(readout["meanCd"], 3) if readout else 0.0
    cl = _fixed(readout["meanCl"], 3) if readout else 0.0
    t, lo, hi = benchmark_target(preset_id, re)
    return BenchmarkResult(preset_id, cd, cl, t, lo, hi, _fixed(abs(cd - t) / t * 100, 1), lo <= cd <= hi, steps, re,
                           cells_across_d)
