"""Python host of the MI355X D3Q19 wind-tunnel solver, over the C ABI of include/aerolbm.h (ctypes).

It mirrors the reference's host-side interface for this path (necrolingus/aerosim):
  * voxelize_mesh_to_lattice  <- voxelizeMeshToLattice (src/geometry/voxelizer.ts:86), run on the GPU
  * VoxelGrid                 <- VoxelGridResult (src/geometry/voxelizer.ts:19-59)
  * LBMSolver                 <- LBMCpuReferenceSolver's object surface (src/solver/lbmCpuReference.ts:55-104):
                                 step(n), reset_equilibrium(), compute_fluid_mass_stats(), f_src, ux/uy/uz/rho
  * LBMEngine                 <- LBMEngine (src/solver/lbmEngine.ts:39-710): set_voxel_grid, update_config,
                                 step_and_extract_slice, reset_flow, get_step_count
  * ForceConvergenceTracker   <- src/solver/forceIntegrator.ts:321-424
The GPU path is the only path: when the HIP library is missing or no device is present the calls raise.
"""
import ctypes
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from .export import NEWTONS_TO_KGF, NEWTONS_TO_LBF, js_to_fixed

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaerolbm.so")

AERO_OK = 0
# error codes (include/aerolbm.h)
AERO_E_ARG = -1      # invalid argument or size
AERO_E_HIP = -2      # HIP runtime error
AERO_E_STATE = -3    # call not valid in the current state
AERO_E_NOMEM = -4    # allocation failed
AERO_E_FORMAT = -5   # input bytes not in the expected file format
AERO_STORE_FP32 = 0
AERO_STORE_FP16 = 1
AERO_PROFILE_UNIFORM = 0
AERO_PROFILE_LOGLAW = 2
AERO_CLOSURE_SMAGORINSKY = 0
AERO_CLOSURE_MLP = 1
AERO_STEP_READOUT = 1
AERO_STEP_MACROS = 2

CELL_FLUID, CELL_SOLID, CELL_BOUNDARY = 0, 1, 2

# lbmConstants.ts:10-63
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
W64 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
W32 = W64.astype(np.float32)


class AeroError(RuntimeError):
    """A library call failed; `code` is its AERO_E_* return code."""

    def __init__(self, message, code=None):
        super().__init__(message)
        self.code = code


class _Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("device", ctypes.c_int32)]


class _Tunnel(ctypes.Structure):
    _fields_ = [("length", ctypes.c_double), ("width", ctypes.c_double), ("height", ctypes.c_double)]


class _Params(ctypes.Structure):
    _fields_ = [("u_inlet", ctypes.c_double), ("nu", ctypes.c_double), ("smagorinsky_cs", ctypes.c_double),
                ("ground_noslip", ctypes.c_int32), ("inflow_profile", ctypes.c_int32),
                ("z0", ctypes.c_double), ("z_ref", ctypes.c_double),
                ("ramp_steps", ctypes.c_int32), ("closure", ctypes.c_int32),
                ("fp16_shift", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("u_init", ctypes.c_double)]


class _ProbeParams(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_double * 3), ("dx", ctypes.c_double), ("wind_speed_mps", ctypes.c_double),
                ("u_inlet_lattice", ctypes.c_double), ("fluid_density", ctypes.c_double)]


class _Refs(ctypes.Structure):
    _fields_ = [("frontal_area", ctypes.c_double), ("char_length", ctypes.c_double),
                ("tunnel_width", ctypes.c_double), ("tunnel_height", ctypes.c_double),
                ("dx", ctypes.c_double), ("u_ref", ctypes.c_double)]


class _Forces(ctypes.Structure):
    _fields_ = [("fx", ctypes.c_double), ("fy", ctypes.c_double), ("fz", ctypes.c_double),
                ("mx", ctypes.c_double), ("my", ctypes.c_double), ("mz", ctypes.c_double),
                ("cut_face_links", ctypes.c_int64),
                ("cd_raw", ctypes.c_double), ("cd", ctypes.c_double), ("cl", ctypes.c_double),
                ("cm", ctypes.c_double), ("xcp", ctypes.c_double)]


_lib = None

_P = ctypes.c_void_p
_SIGS = {
    "aero_create": [ctypes.POINTER(_Grid), ctypes.POINTER(_P)],
    "aero_destroy": [_P],
    "aero_last_error": [_P],
    "aero_set_voxels": [_P, _P, _P, _P, _P, ctypes.c_int32],
    "aero_set_mesh": [_P, _P, ctypes.c_int32, _P, ctypes.POINTER(_Tunnel)],
    "aero_read_voxels": [_P, _P, _P, _P, _P, _P, _P],
    "aero_voxelize_ms": [_P, ctypes.POINTER(ctypes.c_double)],
    "aero_set_params": [_P, ctypes.POINTER(_Params)],
    "aero_reset": [_P, ctypes.c_uint64, ctypes.c_double],
    "aero_step": [_P, ctypes.c_int32, ctypes.c_int32],
    "aero_sync": [_P],
    "aero_step_index": [_P, ctypes.POINTER(ctypes.c_int64)],
    "aero_time_steps": [_P, ctypes.c_int32, _P],
    "aero_set_step_index": [_P, ctypes.c_int64],
    "aero_set_cells_per_thread": [_P, ctypes.c_int32],
    "aero_split_info": [_P, _P],
    "aero_read_forces": [_P, _P, ctypes.POINTER(_Forces)],
    "aero_read_forces_noalloc": [_P, _P, ctypes.POINTER(_Forces)],
    "aero_read_macro": [_P, _P, _P, _P, _P],
    "aero_read_populations": [_P, _P],
    "aero_write_populations": [_P, _P],
    "aero_mass_stats": [_P, _P],
    "aero_anchor_count": [_P, ctypes.POINTER(ctypes.c_int64)],
    "aero_batch_create": [_P, ctypes.c_int32, ctypes.POINTER(_P)],
    "aero_batch_destroy": [_P],
    "aero_batch_last_error": [_P],
    "aero_batch_set_shape": [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32],
    "aero_batch_get_shape": [_P, _P],
    "aero_batch_tune": [_P, ctypes.c_int32, _P, ctypes.POINTER(ctypes.c_double)],
    "aero_batch_step": [_P, ctypes.c_int32],
    "aero_batch_time_steps": [_P, ctypes.c_int32, _P],
    "aero_mlp_create": [ctypes.c_int32, ctypes.POINTER(_P)],
    "aero_mlp_destroy": [_P],
    "aero_mlp_last_error": [_P],
    "aero_mlp_set_params": [_P, _P],
    "aero_mlp_get_params": [_P, _P],
    "aero_mlp_add_samples": [_P, _P, _P, ctypes.c_int64],
    "aero_mlp_clear_samples": [_P],
    "aero_mlp_num_samples": [_P, ctypes.POINTER(ctypes.c_int64)],
    "aero_mlp_get_samples": [_P, _P, _P],
    "aero_mlp_sample_fine": [_P, _P, ctypes.c_int32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                             ctypes.POINTER(ctypes.c_int64)],
    "aero_mlp_grad": [_P, _P, _P, ctypes.c_int32, _P, ctypes.POINTER(ctypes.c_double)],
    "aero_mlp_train": [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _P],
    "aero_mlp_predict": [_P, _P, ctypes.c_int32, _P],
    "aero_set_closure": [_P, _P],
    "aero_mesh_icosphere": [ctypes.c_int32, ctypes.c_double, _P, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
    "aero_mesh_rocket": [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                         ctypes.c_int32, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
    "aero_mesh_naca": [ctypes.c_double, ctypes.c_double, _P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                       ctypes.c_int32, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
    "aero_mesh_box": [_P, _P, _P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
    "aero_mesh_stats": [_P, ctypes.c_int32, _P, _P],
    "aero_read_slice_rgba": [_P, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _P,
                             _P],
    "aero_read_qcriterion": [_P, ctypes.c_double, _P, _P, _P, _P],
    "aero_qcriterion": [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _P, _P, _P, _P,
                        _P, _P, _P, _P],
    "aero_probe": [_P, ctypes.POINTER(_ProbeParams), _P, _P, ctypes.c_int32, _P],
    "aero_stl_parse": [_P, ctypes.c_int64, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)],
    "aero_stl_normalize": [_P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)],
}


def lib():
    """Loads libaerolbm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AeroError(f"{LIB_PATH} is missing: build it with `python -m aerosim_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_char_p if name.endswith("last_error") else ctypes.c_int
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _check(rc, ctx=None, what=""):
    if rc != AERO_OK:
        msg = lib().aero_last_error(ctx)
        raise AeroError(f"{what} failed ({rc}): {msg.decode() if msg else ''}", rc)


# ---------------------------------------------------------------------------------------------
# mesh generators (host C++, no GPU needed)

def _gen(fn, *args):
    n = ctypes.c_int32(0)
    _check(fn(*args, None, 0, ctypes.byref(n)), None, fn.__name__)
    out = np.zeros(n.value * 9, dtype=np.float32)
    _check(fn(*args, _ptr(out), n.value, ctypes.byref(n)), None, fn.__name__)
    return out.reshape(-1, 9)


def mesh_icosphere(level, radius, centre):
    c = np.asarray(centre, dtype=np.float64)
    return _gen(lib().aero_mesh_icosphere, level, radius, _ptr(c))


def mesh_rocket(tip, body_len, diam, nose_len, jitter=0.0, seed=0, segments=128):
    t = np.asarray(tip, dtype=np.float64)
    return _gen(lib().aero_mesh_rocket, _ptr(t), body_len, diam, nose_len, jitter, seed, segments)


def mesh_naca(t, chord, le, alpha_deg, zmin, zmax, npts=64):
    p = np.asarray(le, dtype=np.float64)
    return _gen(lib().aero_mesh_naca, t, chord, _ptr(p), alpha_deg, zmin, zmax, npts)


def mesh_box(lo, hi):
    a = np.asarray(lo, dtype=np.float64)
    b = np.asarray(hi, dtype=np.float64)
    return _gen(lib().aero_mesh_box, _ptr(a), _ptr(b))


def mesh_stats(tris, world=None):
    """computeMeshStats subset (stlLoader.ts:133-244): frontal area, bbox size, volume, open edges."""
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
    w = None if world is None else np.ascontiguousarray(world, dtype=np.float64)
    out = np.zeros(6)
    _check(lib().aero_mesh_stats(_ptr(t), t.size // 9, _ptr(w), _ptr(out)), None, "aero_mesh_stats")
    return {"frontal_area": out[0], "bbox": (out[1], out[2], out[3]), "characteristic_length": max(1e-4, out[1]),
            "volume": out[4], "open_edges": int(out[5]), "watertight": int(out[5]) == 0}


def compute_q_criterion(nx, ny, nz, dx, ux, uy, uz, cell_flags=None, device=0):
    """computeQCriterion (vortexIso.ts:22-119) on the GPU for caller-supplied fields (aero_qcriterion)."""
    N = nx * ny * nz
    a = [np.ascontiguousarray(v, dtype=np.float32).reshape(-1) for v in (ux, uy, uz)]
    if any(v.size != N for v in a):
        raise AeroError("compute_q_criterion: fields need nx*ny*nz values")
    fl = None if cell_flags is None else np.ascontiguousarray(cell_flags, dtype=np.uint8).reshape(-1)
    q = np.zeros(N, dtype=np.float32)
    om = np.zeros(N, dtype=np.float32)
    ov = np.zeros((N, 3), dtype=np.float32)
    mx = np.zeros(2)
    _check(lib().aero_qcriterion(device, nx, ny, nz, dx, *[_ptr(v) for v in a], _ptr(fl), _ptr(q), _ptr(om), _ptr(ov),
                                 _ptr(mx)), None, "aero_qcriterion")
    return {"q_values": q, "omega_mag": om, "omega_vec": ov, "max_q": float(mx[0]), "max_omega": float(mx[1])}


def parse_stl(data):
    """parseSTL (stlLoader.ts:9-95): binary or ASCII STL bytes -> (n, 9) float32 triangle positions.
    Raises AeroError with the reference's message when the bytes are not an STL file."""
    buf = np.frombuffer(bytes(data), dtype=np.uint8)
    n = ctypes.c_int64(0)
    p = _ptr(buf) if buf.size else None
    _check(lib().aero_stl_parse(p, buf.size, None, 0, ctypes.byref(n)), None, "aero_stl_parse")
    out = np.zeros((n.value, 9), dtype=np.float32)
    if n.value:
        _check(lib().aero_stl_parse(p, buf.size, _ptr(out), n.value, ctypes.byref(n)), None, "aero_stl_parse")
    return out


def normalize_and_center_geometry(tris):
    """normalizeAndCenterGeometry (stlLoader.ts:101-127): returns (centred and possibly rescaled copy,
    auto_scaled_from_mm)."""
    t = np.array(tris, dtype=np.float32).reshape(-1, 9)
    flag = ctypes.c_int32(0)
    _check(lib().aero_stl_normalize(_ptr(t) if t.size else None, t.shape[0], ctypes.byref(flag)), None,
           "aero_stl_normalize")
    return t, bool(flag.value)


def load_stl(path):
    """Reads an STL file and applies parse_stl then normalize_and_center_geometry, as the reference's upload
    handler does (main.ts:440-443). Returns (triangles, auto_scaled_from_mm)."""
    with open(path, "rb") as fh:
        return normalize_and_center_geometry(parse_stl(fh.read()))


# ---------------------------------------------------------------------------------------------

@dataclass
class VoxelGrid:
    """VoxelGridResult (voxelizer.ts:19-59)."""
    nx: int
    ny: int
    nz: int
    dx: float
    origin: tuple
    cell_flags: np.ndarray
    boundary_index_map: np.ndarray
    boundary_cell_indices: np.ndarray
    boundary_q_values: np.ndarray
    boundary_min_q: np.ndarray = None
    solid_cell_count: int = 0
    boundary_cell_count: int = 0
    fluid_cell_count: int = 0
    voxelization_time_ms: float = 0.0

    @property
    def dy(self):
        return self.dx

    @property
    def dz(self):
        return self.dx


@dataclass
class Tunnel:
    """TunnelDimensions (types.ts:73-77)."""
    length: float
    width: float
    height: float


@dataclass
class SolverParams:
    u_inlet: float = 0.08
    nu: float = 0.0005
    smagorinsky_cs: float = 0.12
    ground_noslip: bool = False
    inflow_profile: int = AERO_PROFILE_UNIFORM
    z0: float = 0.5
    z_ref: float = 64.0
    ramp_steps: int = 0
    closure: int = AERO_CLOSURE_SMAGORINSKY
    fp16_shift: int = 0
    u_init: float = -1.0

    def to_c(self):
        return _Params(self.u_inlet, self.nu, self.smagorinsky_cs, int(self.ground_noslip), self.inflow_profile,
                       self.z0, self.z_ref, self.ramp_steps, self.closure, self.fp16_shift, 0, self.u_init)


class Context:
    """Owns one aero_ctx (one grid on one device)."""

    def __init__(self, nx, ny, nz, storage=AERO_STORE_FP32, device=0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.N = nx * ny * nz
        self.storage = storage
        h = ctypes.c_void_p()
        g = _Grid(nx, ny, nz, storage, device)
        _check(lib().aero_create(ctypes.byref(g), ctypes.byref(h)), None, "aero_create")
        self.h = h

    def close(self):
        if self.h:
            lib().aero_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        _check(rc, self.h, what)

    def set_voxels(self, vox):
        flags = np.ascontiguousarray(vox.cell_flags, dtype=np.uint8)
        bmap = np.ascontiguousarray(vox.boundary_index_map, dtype=np.int32)
        blist = np.ascontiguousarray(vox.boundary_cell_indices, dtype=np.int32)
        q = np.ascontiguousarray(vox.boundary_q_values, dtype=np.float32)
        if flags.size != self.N or bmap.size != self.N or q.size != 18 * blist.size:
            raise AeroError("set_voxels: cell_flags and boundary_index_map need one entry per cell, "
                            "boundary_q_values 18 per boundary cell", AERO_E_ARG)
        self._c(lib().aero_set_voxels(self.h, _ptr(flags), _ptr(bmap), _ptr(blist), _ptr(q), blist.size),
                "aero_set_voxels")

    def set_mesh(self, tris, tunnel, world=None):
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
        w = None if world is None else np.ascontiguousarray(world, dtype=np.float64)
        tn = _Tunnel(tunnel.length, tunnel.width, tunnel.height)
        self._c(lib().aero_set_mesh(self.h, _ptr(t), t.size // 9, _ptr(w), ctypes.byref(tn)), "aero_set_mesh")

    def read_voxels(self, tunnel_height=None):
        counts = np.zeros(3, dtype=np.int32)
        self._c(lib().aero_read_voxels(self.h, None, None, None, None, None, _ptr(counts)), "aero_read_voxels")
        nb = int(counts[1])
        flags = np.zeros(self.N, dtype=np.uint8)
        bmap = np.zeros(self.N, dtype=np.int32)
        blist = np.zeros(max(nb, 1), dtype=np.int32)
        q = np.zeros(max(nb, 1) * 18, dtype=np.float32)
        minq = np.zeros(max(nb, 1), dtype=np.float32)
        self._c(lib().aero_read_voxels(self.h, _ptr(flags), _ptr(bmap), _ptr(blist), _ptr(q), _ptr(minq),
                                       _ptr(counts)), "aero_read_voxels")
        ms = ctypes.c_double(0)
        lib().aero_voxelize_ms(self.h, ctypes.byref(ms))
        h = (tunnel_height / self.ny) if tunnel_height else 1.0
        return VoxelGrid(self.nx, self.ny, self.nz, h,
                         (-self.nx * h * 0.5, -self.ny * h * 0.5, -self.nz * h * 0.5),
                         flags, bmap, blist[:nb], q[:nb * 18], minq[:nb], int(counts[0]), nb, int(counts[2]),
                         ms.value)

    def set_params(self, p):
        cp = p.to_c()
        self._c(lib().aero_set_params(self.h, ctypes.byref(cp)), "aero_set_params")

    def reset(self, seed=0, noise_amp=0.0):
        self._c(lib().aero_reset(self.h, seed, noise_amp), "aero_reset")

    def step(self, n, readout=False, macros=False):
        """n steps; readout=True: the last one also writes the macro fields and the population snapshot,
        macros=True: only the macro fields (what slices, Q-criterion, probes and mass statistics read)."""
        flags = AERO_STEP_READOUT if readout else (AERO_STEP_MACROS if macros else 0)
        self._c(lib().aero_step(self.h, n, flags), "aero_step")

    def sync(self):
        self._c(lib().aero_sync(self.h), "aero_sync")

    def step_index(self):
        v = ctypes.c_int64(0)
        self._c(lib().aero_step_index(self.h, ctypes.byref(v)), "aero_step_index")
        return v.value

    def set_closure(self, mlp):
        """Uses the Closure `mlp` (None: Smagorinsky). The context keeps a reference: the library reads the
        network's weights on every step, so they must outlive the context's use of them."""
        self._c(lib().aero_set_closure(self.h, mlp.h if mlp is not None else None), "aero_set_closure")
        self._closure = mlp

    def set_cells_per_thread(self, v):
        self._c(lib().aero_set_cells_per_thread(self.h, v), "aero_set_cells_per_thread")

    def set_step_index(self, step):
        self._c(lib().aero_set_step_index(self.h, step), "aero_set_step_index")

    def anchor_count(self):
        v = ctypes.c_int64(0)
        self._c(lib().aero_anchor_count(self.h, ctypes.byref(v)), "aero_anchor_count")
        return v.value

    def split_info(self):
        """Cells updated by the bulk kernel, cells visited by the special-cell kernel, cells per thread."""
        out = np.zeros(4, dtype=np.int64)
        self._c(lib().aero_split_info(self.h, _ptr(out)), "aero_split_info")
        return {"bulk_cells": int(out[0]), "special_cells": int(out[1]), "cells_per_thread": int(out[2]),
                "face_cells": int(out[3])}

    def time_steps(self, n):
        out = np.zeros(5)
        self._c(lib().aero_time_steps(self.h, n, _ptr(out)), "aero_time_steps")
        return {"total_ms": out[0], "kernel_ms": out[1], "step_ms": out[2], "bulk_ms": out[3], "after_bulk_ms": out[4]}

    def read_forces(self, refs=None, noalloc=False):
        """Force, torque and coefficients of the current state. noalloc=True calls aero_read_forces_noalloc, the
        variant that allocates nothing (fixed per-GPU scratch; up to 2^20 boundary cells)."""
        f = _Forces()
        r = None
        if refs is not None:
            r = _Refs(refs["frontal_area"], refs["char_length"], refs["tunnel_width"], refs["tunnel_height"],
                      refs["dx"], refs["u_ref"])
        fn = lib().aero_read_forces_noalloc if noalloc else lib().aero_read_forces
        name = "aero_read_forces_noalloc" if noalloc else "aero_read_forces"
        self._c(fn(self.h, ctypes.byref(r) if r is not None else None, ctypes.byref(f)), name)
        return {k: getattr(f, k) for k, _ in _Forces._fields_}

    def read_macro(self):
        out = [np.zeros(self.N, dtype=np.float32) for _ in range(4)]
        self._c(lib().aero_read_macro(self.h, *[_ptr(a) for a in out]), "aero_read_macro")
        return out

    SLICE_AXES = {"z": 0, "y": 1, "x": 2}
    FIELD_MODES = {"velocity": 0, "vorticity": 1, "cp": 2}

    def extract_slice_rgba(self, axis="z", slice_fraction=0.5, field_mode="vorticity", rules="cpu", u_ref=None):
        """extractSliceRGBA (lbmCpuReference.ts:500-596) on the GPU: (width, height, rgba uint8 array of
        height x width x 4). rules="wgsl" applies the post-process shader's rules (postProcess.wgsl:52-146).
        Needs the macro fields of the last step (step with readout=True)."""
        wh = np.zeros(2, dtype=np.int32)
        a, f = self.SLICE_AXES[axis], self.FIELD_MODES[field_mode]
        r = {"cpu": 0, "wgsl": 1}[rules]
        ur = -1.0 if u_ref is None else float(u_ref)
        self._c(lib().aero_read_slice_rgba(self.h, a, slice_fraction, f, r, ur, None, _ptr(wh)), "aero_read_slice_rgba")
        out = np.zeros((int(wh[1]), int(wh[0]), 4), dtype=np.uint8)
        self._c(lib().aero_read_slice_rgba(self.h, a, slice_fraction, f, r, ur, _ptr(out), _ptr(wh)),
                "aero_read_slice_rgba")
        return int(wh[0]), int(wh[1]), out

    def compute_q_criterion(self, dx=1.0, fields=True):
        """computeQCriterion (vortexIso.ts:22-119) on the GPU over the current velocity field and cell flags:
        dict of q_values, omega_mag, omega_vec (N x 3), max_q, max_omega. fields=False returns the maxima only
        (the fields stay on the device)."""
        q = np.zeros(self.N, dtype=np.float32) if fields else None
        om = np.zeros(self.N, dtype=np.float32) if fields else None
        ov = np.zeros((self.N, 3), dtype=np.float32) if fields else None
        mx = np.zeros(2)
        self._c(lib().aero_read_qcriterion(self.h, dx, _ptr(q), _ptr(om), _ptr(ov), _ptr(mx)), "aero_read_qcriterion")
        return {"q_values": q, "omega_mag": om, "omega_vec": ov, "max_q": float(mx[0]), "max_omega": float(mx[1])}

    PROBE_FIELDS = ("ix", "iy", "iz", "speed_mps", "speed_mph", "ux_mps", "uy_mps", "uz_mps", "gauge_pressure_pa",
                    "pressure_coeff_cp", "density_rho", "dynamic_pressure_pa", "is_inside_solid")

    def probe(self, points, origin, dx, wind_speed_mps=25.0, u_inlet_lattice=0.085, fluid_density=1.204,
              on_surface=None):
        """Point probe readouts (pointProbe.ts:284-355 and computeProbePhysics, :23-92) at world points
        (n x 3): an (n x 13) array whose columns are PROBE_FIELDS."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        n = pts.shape[0]
        surf = None if on_surface is None else np.ascontiguousarray(on_surface, dtype=np.uint8).reshape(-1)
        if surf is not None and surf.size != n:
            raise AeroError("probe: on_surface needs one flag per point")
        pp = _ProbeParams((ctypes.c_double * 3)(*[float(v) for v in origin]), dx, wind_speed_mps, u_inlet_lattice,
                          fluid_density)
        out = np.zeros((n, 13))
        self._c(lib().aero_probe(self.h, ctypes.byref(pp), _ptr(pts), _ptr(surf), n, _ptr(out)), "aero_probe")
        return out

    def read_populations(self):
        f = np.zeros(19 * self.N, dtype=np.float32)
        self._c(lib().aero_read_populations(self.h, _ptr(f)), "aero_read_populations")
        return f

    def write_populations(self, f):
        a = np.ascontiguousarray(f, dtype=np.float32).reshape(-1)
        if a.size != 19 * self.N:
            raise AeroError("write_populations: need 19*N values")
        self._c(lib().aero_write_populations(self.h, _ptr(a)), "aero_write_populations")

    def mass_stats(self):
        out = np.zeros(5)
        self._c(lib().aero_mass_stats(self.h, _ptr(out)), "aero_mass_stats")
        return {"total_mass": out[0], "mean_rho": out[1], "max_speed": out[2], "has_nan": bool(out[3]),
                "fluid_count": int(out[4])}


class Batch:
    """Many contexts stepped by one persistent, graph-captured launch (aero_batch_*)."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = ctypes.c_void_p()
        rc = lib().aero_batch_create(arr, len(self.contexts), ctypes.byref(h))
        if rc != AERO_OK:
            raise AeroError(f"aero_batch_create failed ({rc}): {lib().aero_batch_last_error(None).decode()}", rc)
        self.h = h

    def _c(self, rc, what):
        if rc != AERO_OK:
            raise AeroError(f"{what} failed ({rc}): {lib().aero_batch_last_error(self.h).decode()}", rc)

    def close(self):
        if self.h:
            lib().aero_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, block, cpt, grid):
        self._c(lib().aero_batch_set_shape(self.h, block, cpt, grid), "aero_batch_set_shape")

    def shape(self):
        out = np.zeros(3, dtype=np.int32)
        self._c(lib().aero_batch_get_shape(self.h, _ptr(out)), "aero_batch_get_shape")
        return {"block": int(out[0]), "cells_per_thread": int(out[1]), "grid": int(out[2])}

    def tune(self, steps=4):
        out = np.zeros(3, dtype=np.int32)
        ms = ctypes.c_double(0)
        self._c(lib().aero_batch_tune(self.h, steps, _ptr(out), ctypes.byref(ms)), "aero_batch_tune")
        return {"block": int(out[0]), "cells_per_thread": int(out[1]), "grid": int(out[2]), "ms_per_step": ms.value}

    def step(self, n):
        self._c(lib().aero_batch_step(self.h, n), "aero_batch_step")

    def time_steps(self, n):
        out = np.zeros(2)
        self._c(lib().aero_batch_time_steps(self.h, n, _ptr(out)), "aero_batch_time_steps")
        return {"total_ms": out[0], "step_ms": out[1]}


MLP_INPUTS, MLP_HIDDEN = 6, 32
MLP_NPARAM = MLP_HIDDEN * MLP_INPUTS + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + 2 * MLP_HIDDEN + 1  # 1313


class Closure:
    """Learned eddy-viscosity closure (aero_mlp_*): a 6 -> 32 -> 32 -> 1 MLP trained on the GPU."""

    def __init__(self, device=0, params=None):
        h = ctypes.c_void_p()
        rc = lib().aero_mlp_create(device, ctypes.byref(h))
        if rc != AERO_OK:
            raise AeroError(f"aero_mlp_create failed ({rc})", rc)
        self.h = h
        if params is not None:
            self.set_params(params)

    def _c(self, rc, what):
        if rc != AERO_OK:
            raise AeroError(f"{what} failed ({rc}): {lib().aero_mlp_last_error(self.h).decode()}", rc)

    def close(self):
        if self.h:
            lib().aero_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != MLP_NPARAM:
            raise AeroError("closure parameters: need 1313 values")
        self._c(lib().aero_mlp_set_params(self.h, _ptr(a)), "aero_mlp_set_params")

    def params(self):
        a = np.zeros(MLP_NPARAM, dtype=np.float32)
        self._c(lib().aero_mlp_get_params(self.h, _ptr(a)), "aero_mlp_get_params")
        return a

    def add_samples(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1)
        self._c(lib().aero_mlp_add_samples(self.h, _ptr(X), _ptr(T), T.size), "aero_mlp_add_samples")

    def clear_samples(self):
        self._c(lib().aero_mlp_clear_samples(self.h), "aero_mlp_clear_samples")

    def num_samples(self):
        n = ctypes.c_int64(0)
        self._c(lib().aero_mlp_num_samples(self.h, ctypes.byref(n)), "aero_mlp_num_samples")
        return n.value

    def samples(self):
        n = self.num_samples()
        X = np.zeros((n, 8), dtype=np.float32)
        T = np.zeros(n, dtype=np.float32)
        self._c(lib().aero_mlp_get_samples(self.h, _ptr(X), _ptr(T)), "aero_mlp_get_samples")
        return X, T

    def sample_fine(self, fine_ctx, stride, seed, nu_coarse, cs=0.17):
        n = ctypes.c_int64(0)
        self._c(lib().aero_mlp_sample_fine(self.h, fine_ctx.h, stride, seed, nu_coarse, cs, ctypes.byref(n)),
                "aero_mlp_sample_fine")
        return n.value

    def grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
        T = np.ascontiguousarray(T, dtype=np.float32).reshape(-1)
        g = np.zeros(MLP_NPARAM, dtype=np.float32)
        loss = ctypes.c_double(0)
        self._c(lib().aero_mlp_grad(self.h, _ptr(X), _ptr(T), T.size, _ptr(g), ctypes.byref(loss)), "aero_mlp_grad")
        return loss.value, g

    def train(self, iters, batch=4096, lr=1e-3):
        losses = np.zeros(max(iters, 1), dtype=np.float32)
        self._c(lib().aero_mlp_train(self.h, iters, batch, lr, _ptr(losses)), "aero_mlp_train")
        return losses[:iters]

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 8)
        y = np.zeros(X.shape[0], dtype=np.float32)
        self._c(lib().aero_mlp_predict(self.h, _ptr(X), X.shape[0], _ptr(y)), "aero_mlp_predict")
        return y


def voxelize_mesh_to_lattice(tris, world, tunnel, grid, device=0):
    """voxelizeMeshToLattice (voxelizer.ts:86) on the GPU; returns a VoxelGrid."""
    nx, ny, nz = grid
    ctx = Context(nx, ny, nz, device=device)
    try:
        ctx.set_mesh(tris, tunnel, world)
        return ctx.read_voxels(tunnel.height)
    finally:
        ctx.close()


class LBMSolver:
    """GPU solver with the object surface of LBMCpuReferenceSolver (lbmCpuReference.ts:55-104)."""

    def __init__(self, voxel, u_inlet=0.08, nu_lattice=0.0005, smagorinsky_cs=0.12, ground_noslip=False,
                 storage=AERO_STORE_FP32, device=0, seed=0, noise_amp=0.0, **extra):
        self.voxel = voxel
        self.nx, self.ny, self.nz = voxel.nx, voxel.ny, voxel.nz
        self.total_cells = self.nx * self.ny * self.nz
        self.ctx = Context(self.nx, self.ny, self.nz, storage=storage, device=device)
        self.ctx.set_voxels(voxel)
        self.params = SolverParams(u_inlet=u_inlet, nu=nu_lattice, smagorinsky_cs=smagorinsky_cs,
                                   ground_noslip=ground_noslip, **extra)
        self.seed, self.noise_amp = seed, noise_amp
        self._fields = None
        self.reset_equilibrium()

    # parameters are plain attributes in the reference; keep them in sync with the device on step()
    @property
    def u_inlet(self):
        return self.params.u_inlet

    @u_inlet.setter
    def u_inlet(self, v):
        self.params.u_inlet = float(v)
        self.ctx.set_params(self.params)

    def reset_equilibrium(self):
        self.ctx.set_params(self.params)
        self.ctx.reset(self.seed, self.noise_amp)
        self._fields = None

    def step(self, num_steps=1, readout=True):
        self.ctx.step(num_steps, readout=readout)
        self._fields = None

    @property
    def step_index(self):
        return self.ctx.step_index()

    def _read(self):
        if self._fields is None:
            ux, uy, uz, rho = self.ctx.read_macro()
            self._fields = {"ux": ux, "uy": uy, "uz": uz, "rho": rho}
        return self._fields

    @property
    def ux(self):
        return self._read()["ux"]

    @property
    def uy(self):
        return self._read()["uy"]

    @property
    def uz(self):
        return self._read()["uz"]

    @property
    def rho(self):
        return self._read()["rho"]

    @property
    def f_src(self):
        return self.ctx.read_populations()

    def compute_fluid_mass_stats(self):
        s = self.ctx.mass_stats()
        return {"totalMass": s["total_mass"], "meanRho": s["mean_rho"], "maxSpeed": s["max_speed"],
                "hasNaN": s["has_nan"]}

    def forces(self, refs=None):
        return self.ctx.read_forces(refs)

    def extract_slice_rgba(self, axis="z", slice_fraction=0.5, field_mode="velocity"):
        """extractSliceRGBA (lbmCpuReference.ts:500-596) on the GPU, uRef = max(1e-4, u_inlet): dict of width,
        height and rgba (flat uint8, 4 per pixel, row-major)."""
        w, h, img = self.ctx.extract_slice_rgba(axis, slice_fraction, field_mode, "cpu", self.params.u_inlet)
        return {"width": w, "height": h, "rgba": img.reshape(-1)}

    def close(self):
        self.ctx.close()


class ForceConvergenceTracker:
    """forceIntegrator.ts:321-424 (rolling window of Cd/Cl/Cm/Xcp samples, history capped at 180)."""

    def __init__(self, window_size=60):
        self.reset(window_size)

    def reset(self, window_size=60):
        self.window_size = window_size
        self.samples = []

    def push_sample(self, step, cd, cl, cm, xcp, wind_speed=25.0, fluid_density=1.204, frontal_area_m2=1.0):
        self.samples.append({"step": step, "cd": cd, "cl": cl, "cm": cm, "xcp": xcp})
        if len(self.samples) > 180:
            self.samples.pop(0)
        recent = self.samples[-self.window_size:]
        n = len(recent)
        sums = {"cd": 0.0, "cl": 0.0, "cm": 0.0, "xcp": 0.0}
        for r in recent:  # in sample order, as the reference sums
            for k in sums:
                sums[k] += r[k]
        mean = {k: v / n for k, v in sums.items()}
        var = 0.0
        for r in recent:
            d = r["cd"] - mean["cd"]
            var += d * d
        std = math.sqrt(var / max(1, n))
        rel = std / max(0.02, abs(mean["cd"]))
        fill = min(1.0, n / 25)
        stability = float(js_to_fixed(max(0.0, min(99.9, fill * (1.0 - min(1.0, rel * 4.0)) * 100)), 1))
        q_area = 0.5 * fluid_density * wind_speed * wind_speed * frontal_area_m2
        drag = q_area * mean["cd"]
        lift = q_area * mean["cl"]
        return {"instantCd": cd, "instantCl": cl, "meanCd": mean["cd"], "meanCl": mean["cl"], "stdCd": std,
                "dragForceNewtons": drag, "dragForceLbs": drag * NEWTONS_TO_LBF, "dragForceKgf": drag * NEWTONS_TO_KGF,
                "liftForceNewtons": lift, "liftForceLbs": lift * NEWTONS_TO_LBF,
                "pitchingMomentCm": mean["cm"], "centerOfPressureX": mean["xcp"], "stabilityPercent": stability,
                "isConverged": n >= 20 and stability >= 96.0,
                "history": [{"step": r["step"], "cd": r["cd"], "cl": r["cl"], "cm": r["cm"], "xcp": r["xcp"],
                             "dragForceN": r["cd"] * q_area, "liftForceN": r["cl"] * q_area} for r in self.samples]}


@dataclass
class SolverConfig:
    """Numerical and reference fields of LBMSolverConfig (lbmEngine.ts:17-32)."""
    u_inlet_lattice: float = 0.085
    nu_lattice: float = 0.0005
    smagorinsky_cs: float = 0.12
    ground_noslip: bool = False
    wind_speed_mps: float = 25.0
    fluid_density: float = 1.204
    fluid_kinematic_viscosity: float = 1.516e-5   # air_20c, types.ts:16-22
    fluid_speed_of_sound: float = 343.2
    frontal_area_m2: float = math.pi * 0.25
    characteristic_length_m: float = 1.0
    tunnel: Tunnel = field(default_factory=lambda: Tunnel(6.0, 2.8, 2.4))
    slice_axis: str = "z"          # lbmEngine.ts:47-49
    slice_fraction: float = 0.5
    field_mode: str = "vorticity"
    slice_rules: str = "wgsl"      # the GPU engine's post-process shader; "cpu" for its CPU fallback's rules


class LBMEngine:
    """LBMEngine (lbmEngine.ts:39-710) on the HIP path: geometry, config, substeps, force readout."""

    def __init__(self, device=0, storage=AERO_STORE_FP32):
        self.device, self.storage = device, storage
        self.voxel = None
        self.ctx = None
        self.config = SolverConfig()
        self.tracker = ForceConvergenceTracker(60)
        self.step_count = 0
        self.latest_readout = None

    def set_voxel_grid(self, voxel, config=None, reset_flow=True):
        if config:
            self.update_config(**config)
        if self.ctx is None or (self.ctx.nx, self.ctx.ny, self.ctx.nz) != (voxel.nx, voxel.ny, voxel.nz):
            if self.ctx is not None:
                self.ctx.close()
            self.ctx = Context(voxel.nx, voxel.ny, voxel.nz, storage=self.storage, device=self.device)
        self.voxel = voxel
        self.ctx.set_voxels(voxel)
        self._push_params()
        if reset_flow:
            self.reset_flow()

    def update_config(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.config, k):
                raise AeroError(f"unknown config field {k}")
            setattr(self.config, k, v)
        if self.ctx is not None:
            self._push_params()

    def _push_params(self):
        c = self.config
        self.ctx.set_params(SolverParams(u_inlet=c.u_inlet_lattice, nu=c.nu_lattice, smagorinsky_cs=c.smagorinsky_cs,
                                         ground_noslip=c.ground_noslip))

    def reset_flow(self):
        self.ctx.reset(0, 0.0)
        self.step_count = 0
        self.tracker.reset(60)

    def get_step_count(self):
        return self.step_count

    def refs(self):
        c, v = self.config, self.voxel
        return {"frontal_area": c.frontal_area_m2, "char_length": c.characteristic_length_m,
                "tunnel_width": c.tunnel.width, "tunnel_height": c.tunnel.height, "dx": v.dx,
                "u_ref": c.u_inlet_lattice}

    def step_and_extract_slice(self, substeps=4, read_macro_field=False):
        """stepAndExtractSlice (lbmEngine.ts:420-660): substeps, the slice of the configured axis, fraction and
        field mode, the force readout, and optionally the macro fields."""
        if self.ctx is None:
            return None
        self.ctx.step(substeps, macros=True)
        self.step_count += substeps
        f = self.ctx.read_forces(self.refs())
        c = self.config
        readout = self.tracker.push_sample(self.step_count, f["cd"], f["cl"], f["cm"], f["xcp"], c.wind_speed_mps,
                                           c.fluid_density, c.frontal_area_m2)
        w, h, img = self.ctx.extract_slice_rgba(c.slice_axis, c.slice_fraction, c.field_mode, c.slice_rules,
                                                c.u_inlet_lattice)
        out = {"width": w, "height": h, "rgba": img.reshape(-1), "stepCount": self.step_count,
               "forceReadout": readout, "forces": f}
        self.latest_readout = readout
        if read_macro_field:
            out["macroField"] = self.ctx.read_macro()
        return out

    SUBSTEPS_PER_FRAME = 4  # main.ts:490, with a GPU device

    def tick(self, need_macro=False, substeps=SUBSTEPS_PER_FRAME):
        """One frame of the reference's solver loop (main.ts:491-512): stepAndExtractSlice over the frame's
        substeps, the frame's wall time dt (at least 0.1 ms) and MLUPS = cells * substeps / (dt * 1e6)."""
        t0 = time.perf_counter()
        out = self.step_and_extract_slice(substeps, need_macro)
        if out is None:
            return None
        dt = max(1e-4, time.perf_counter() - t0)
        v = self.voxel
        out["mlups"] = (v.nx * v.ny * v.nz * substeps) / (dt * 1e6)
        out["frameSeconds"] = dt
        return out

    def _export_inputs(self, specimen_name, mesh_stats, timestamp):
        from . import export as X

        r, c = self.latest_readout, self.config
        if r is None:
            raise AeroError("no force readout yet: call step_and_extract_slice or tick first")
        ms = dict(mesh_stats or {"frontalArea": c.frontal_area_m2, "characteristicLength": c.characteristic_length_m})
        fluid = {"density": c.fluid_density, "kinematicViscosity": c.fluid_kinematic_viscosity,
                 "speedOfSound": c.fluid_speed_of_sound}
        tunnel = {"length": c.tunnel.length, "width": c.tunnel.width, "height": c.tunnel.height}
        aero = X.compute_aerodynamic_telemetry(c.wind_speed_mps, ms.get("characteristicLength", c.characteristic_length_m),
                                               ms.get("frontalArea", c.frontal_area_m2), fluid, tunnel)
        hist = X.history_for_export(r, c.wind_speed_mps, c.fluid_density, ms.get("frontalArea", c.frontal_area_m2))
        meta = {"timestamp": timestamp or X.iso_now(), "specimenName": specimen_name, "solverMode": "interactive",
                "groundPlaneEnabled": bool(c.ground_noslip), "windSpeedMps": c.wind_speed_mps, "fluid": fluid,
                "tunnel": tunnel, "meshStats": ms, "meanCd": r["meanCd"], "meanCl": r["meanCl"],
                "meanCm": r["pitchingMomentCm"], "centerOfPressureX": r["centerOfPressureX"],
                "dragForceN": r["dragForceNewtons"], "liftForceN": r["liftForceNewtons"], "reynolds": aero["reynoldsNumber"],
                "dynamicPressurePa": aero["dynamicPressurePa"], "blockageRatioPercent": aero["blockageRatioPercent"],
                "totalSteps": r["history"][-1]["step"] if r["history"] else 0, "isConverged": r["isConverged"]}
        return X, meta, hist

    def export_telemetry_csv(self, specimen_name="Specimen", mesh_stats=None, timestamp=None):
        """The CSV of the reference's export button (main.ts:294-332) from the latest force readout."""
        X, meta, hist = self._export_inputs(specimen_name, mesh_stats, timestamp)
        return X.export_telemetry_to_csv(hist, meta)

    def export_run_report_json(self, specimen_name="Specimen", mesh_stats=None, timestamp=None):
        """The JSON report of the reference's export button (main.ts:333-385) from the latest force readout."""
        X, meta, hist = self._export_inputs(specimen_name, mesh_stats, timestamp)
        return X.export_run_report_to_json(meta, hist, timestamp)
