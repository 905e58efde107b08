"""ctypes wrapper of the CPU oracle (oracle/lbm_oracle.cpp). TEST INFRASTRUCTURE ONLY: loaded by tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py, never by the product package."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None
_P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        sig = {
            "orc_new": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int], _P),
            "orc_free": ([_P], None),
            "orc_set_threads": ([ctypes.c_int], None),
            "orc_max_threads": ([], ctypes.c_int),
            "orc_set_geometry": ([_P, _P, _P, _P, _P, ctypes.c_int], None),
            "orc_set_params": ([_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, ctypes.c_double], None),
            "orc_set_uinlet": ([_P, ctypes.c_double], None),
            "orc_reset": ([_P, ctypes.c_uint64, ctypes.c_double], None),
            "orc_step": ([_P, ctypes.c_int], None),
            "orc_step_index": ([_P], ctypes.c_long),
            "orc_set_anchor": ([_P, ctypes.c_int], None),
            "orc_set_closure": ([_P, _P], None),
            "orc_closure_features": ([_P, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P], None),
            "orc_closure_forward": ([_P, _P, ctypes.c_long, _P], None),
            "orc_anchor_count": ([_P], ctypes.c_int),
            "orc_get_f": ([_P, _P], None),
            "orc_set_f": ([_P, _P], None),
            "orc_get_macro": ([_P, _P, _P, _P, _P], None),
            "orc_forces": ([_P, _P], None),
            "orc_mass_stats": ([_P, _P], None),
            "orc_mass_stats_fields": ([_P, _P, _P, _P, _P, ctypes.c_long, _P], None),
            "orc_vox_new": ([_P, ctypes.c_int, _P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int], _P),
            "orc_vox_counts": ([_P, _P], None),
            "orc_vox_geom": ([_P, _P], None),
            "orc_vox_get": ([_P, _P, _P, _P, _P, _P], None),
            "orc_vox_free": ([_P], None),
            "orc_bvh_new": ([_P, ctypes.c_int, _P], _P),
            "orc_bvh_free": ([_P], None),
            "orc_bvh_inside": ([_P, ctypes.c_double, ctypes.c_double, ctypes.c_double], ctypes.c_int),
            "orc_bvh_segment": ([_P] + [ctypes.c_double] * 7, ctypes.c_double),
            "orc_bvh_scanline": ([_P, ctypes.c_double, ctypes.c_double, _P, ctypes.c_int], ctypes.c_int),
            "orc_frontal_area": ([_P, ctypes.c_int, _P, ctypes.c_int], ctypes.c_double),
            "orc_feq64": ([ctypes.c_int] + [ctypes.c_double] * 4, ctypes.c_double),
            "orc_noise_unit": ([ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int], ctypes.c_double),
        }
        for k, (a, r) in sig.items():
            fn = getattr(L, k)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Voxels:
    """Oracle VoxelGridResult for a triangle soup (voxelizer.ts:86-281)."""

    def __init__(self, tris, tunnel_height, grid, world=None):
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
        w = None if world is None else np.ascontiguousarray(world, dtype=np.float64)
        nx, ny, nz = grid
        h = lib().orc_vox_new(_p(t), t.size // 9, _p(w), float(tunnel_height), nx, ny, nz)
        c = np.zeros(3, dtype=np.int32)
        lib().orc_vox_counts(h, _p(c))
        g = np.zeros(4)
        lib().orc_vox_geom(h, _p(g))
        N = nx * ny * nz
        nb = int(c[1])
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dx, self.origin = g[0], (g[1], g[2], g[3])
        self.cell_flags = np.zeros(N, dtype=np.uint8)
        self.boundary_index_map = np.zeros(N, dtype=np.int32)
        self.boundary_cell_indices = np.zeros(max(nb, 1), dtype=np.int32)
        self.boundary_q_values = np.zeros(max(nb, 1) * 18, dtype=np.float32)
        self.boundary_min_q = np.zeros(max(nb, 1), dtype=np.float32)
        lib().orc_vox_get(h, _p(self.cell_flags), _p(self.boundary_index_map), _p(self.boundary_cell_indices),
                          _p(self.boundary_q_values), _p(self.boundary_min_q))
        lib().orc_vox_free(h)
        self.boundary_cell_indices = self.boundary_cell_indices[:nb]
        self.boundary_q_values = self.boundary_q_values[:nb * 18]
        self.boundary_min_q = self.boundary_min_q[:nb]
        self.solid_cell_count, self.boundary_cell_count, self.fluid_cell_count = int(c[0]), nb, int(c[2])


class BVH:
    def __init__(self, tris, world=None):
        t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
        w = None if world is None else np.ascontiguousarray(world, dtype=np.float64)
        self.h = lib().orc_bvh_new(_p(t), t.size // 9, _p(w))

    def inside(self, x, y, z):
        return bool(lib().orc_bvh_inside(self.h, x, y, z))

    def segment(self, o, d, max_t=1.0):
        return lib().orc_bvh_segment(self.h, *o, *d, max_t)

    def scanline(self, y, z, cap=256):
        out = np.zeros(cap)
        n = lib().orc_bvh_scanline(self.h, y, z, _p(out), cap)
        return out[:min(n, cap)]

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_free(self.h)
            self.h = None


def frontal_area(tris, world=None, grid=512):
    t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1)
    w = None if world is None else np.ascontiguousarray(world, dtype=np.float64)
    return lib().orc_frontal_area(_p(t), t.size // 9, _p(w), grid)


class Solver:
    """Oracle solver: mode 'fp64' (all double) or 'ref32' (reference-faithful float storage)."""

    def __init__(self, vox, mode="fp64", u_inlet=0.08, nu=0.0005, cs=0.12, ground_noslip=False, profile=0, z0=0.5,
                 zref=64.0):
        self.nx, self.ny, self.nz = vox.nx, vox.ny, vox.nz
        self.N = self.nx * self.ny * self.nz
        self.h = lib().orc_new(self.nx, self.ny, self.nz, 1 if mode == "ref32" else 0)
        flags = np.ascontiguousarray(vox.cell_flags, dtype=np.uint8)
        bmap = np.ascontiguousarray(vox.boundary_index_map, dtype=np.int32)
        blist = np.ascontiguousarray(vox.boundary_cell_indices, dtype=np.int32)
        q = np.ascontiguousarray(vox.boundary_q_values, dtype=np.float32)
        lib().orc_set_geometry(self.h, _p(flags), _p(bmap), _p(blist), _p(q), blist.size)
        lib().orc_set_params(self.h, u_inlet, nu, cs, int(ground_noslip), profile, z0, zref)

    def set_u_inlet(self, u):
        lib().orc_set_uinlet(self.h, u)

    def reset(self, seed=0, noise_amp=0.0):
        lib().orc_reset(self.h, seed, noise_amp)

    def step(self, n=1):
        lib().orc_step(self.h, n)

    @property
    def step_index(self):
        return lib().orc_step_index(self.h)

    def set_closure(self, params):
        """Learned closure with these 1313 parameters (the C ABI's order), or None for Smagorinsky."""
        self._closure = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        lib().orc_set_closure(self.h, _p(self._closure))

    def set_anchor(self, on):
        """Switches the global mass anchor (lbmCpuReference.ts:431-455) on or off (tests compare a step with and
        without it)."""
        lib().orc_set_anchor(self.h, int(bool(on)))

    def anchor_count(self):
        return lib().orc_anchor_count(self.h)

    def populations(self):
        out = np.zeros(19 * self.N)
        lib().orc_get_f(self.h, _p(out))
        return out

    def set_populations(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_set_f(self.h, _p(a))

    def macro(self):
        out = [np.zeros(self.N) for _ in range(4)]
        lib().orc_get_macro(self.h, *[_p(a) for a in out])
        return out

    def forces(self):
        out = np.zeros(7)
        lib().orc_forces(self.h, _p(out))
        return {"fx": out[0], "fy": out[1], "fz": out[2], "mx": out[3], "my": out[4], "mz": out[5],
                "cut_face_links": int(out[6])}

    def mass_stats(self):
        out = np.zeros(5)
        lib().orc_mass_stats(self.h, _p(out))
        return {"total_mass": out[0], "mean_rho": out[1], "max_speed": out[2], "has_nan": bool(out[3]),
                "fluid_count": int(out[4])}

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None


def feq64(d, rho, ux, uy, uz):
    return lib().orc_feq64(d, rho, ux, uy, uz)


def noise_unit(seed, idx, comp):
    return lib().orc_noise_unit(seed, idx, comp)


def coefficients(f, frontal_area, char_length, tunnel_width, tunnel_height, dx, u_ref):
    """computeNormalizedForceCoefficients (forceIntegrator.ts:240-315) in double."""
    cell_area = dx * dx
    q_lat = max(1e-9, 0.5 * 1.0 * u_ref * u_ref)
    a_ref = max(1e-4, frontal_area / cell_area)
    l_ref = max(1.0, char_length / dx)
    cd_raw = f["fx"] / (q_lat * a_ref)
    cl = f["fy"] / (q_lat * a_ref)
    cm = f["mz"] / (q_lat * a_ref * l_ref)
    tunnel_area = max(1e-4, tunnel_width * tunnel_height)
    blockage = min(0.40, frontal_area / tunnel_area)
    cd = cd_raw / (1.0 + 2.5 * max(0.0, cd_raw) * blockage)
    xcp = (max(-char_length * 0.5, min(char_length * 0.5, -(f["mz"] / f["fy"]) * dx))
           if abs(f["fy"]) > 1e-5 else 0.0)
    return {"cd_raw": cd_raw, "cd": cd, "cl": cl, "cm": cm, "xcp": xcp}


def mass_stats_fields(flags, rho, ux, uy, uz):
    """computeFluidMassStats (lbmCpuReference.ts:461-495) over given float32 fields."""
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    a = [np.ascontiguousarray(x, dtype=np.float32) for x in (rho, ux, uy, uz)]
    out = np.zeros(5)
    lib().orc_mass_stats_fields(_p(fl), *[_p(x) for x in a], fl.size, _p(out))
    return {"total_mass": out[0], "mean_rho": out[1], "max_speed": out[2], "has_nan": bool(out[3]),
            "fluid_count": int(out[4])}


def set_threads(n):
    """Threads of the oracle's parallel step loop."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def closure_features(P, nu, tau0, smag):
    """The learned closure's 6 features of a non-equilibrium stress per unit density (xx, yy, zz, xy, xz, yz)."""
    a = np.ascontiguousarray(P, dtype=np.float64)
    out = np.zeros(6)
    lib().orc_closure_features(_p(a), nu, tau0, smag, _p(out))
    return out


def closure_forward(params, X):
    """Network output of each row of X (n x 6) for 1313 parameters."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(np.asarray(X, dtype=np.float64)[:, :6])
    out = np.zeros(x.shape[0])
    lib().orc_closure_forward(_p(p), _p(x), x.shape[0], _p(out))
    return out
