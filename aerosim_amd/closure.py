"""Learned-closure workflow on the GPU: training data from filtered 2x-resolution runs that the solver itself
generates, training with the matrix-core kernels, and the judge comparison (coarse learned vs coarse
Smagorinsky against the fine-grid plain run).

Workload (ground boundary layer around a block building): the lattice's vertical axis is y (floor at
iy = 0), as in the reference (voxelizer.ts:101, lbmCpuReference.ts:375-401); a workload quoted as
nx x lateral x height maps to a grid (nx, height, lateral). Inflow is the log law
u(y) = u_ref ln((y + 1/2)/z0) / ln(z_ref/z0) with a no-slip floor.
"""
from dataclasses import dataclass

import numpy as np

from . import lbm as L


@dataclass
class BuildingCase:
    """Coarse-grid description; the fine grid doubles every length."""
    nx: int = 512            # streamwise
    height: int = 128        # vertical (lattice y)
    lateral: int = 256       # spanwise (lattice z)
    bx: float = 40.0         # building streamwise size (cells)
    bz: float = 20.0         # building lateral size
    bh: float = 60.0         # building height
    x_front: float = 0.25    # building front at this fraction of nx
    u_ref: float = 0.08
    z0: float = 0.5
    z_ref: float = 64.0
    re: float = 1e5
    cs: float = 0.17

    def scaled(self, f):
        return BuildingCase(int(self.nx * f), int(self.height * f), int(self.lateral * f), self.bx * f, self.bz * f,
                            self.bh * f, self.x_front, self.u_ref, self.z0 * f, self.z_ref * f, self.re, self.cs)

    @property
    def nu(self):
        return self.u_ref * self.bh / self.re

    def mesh(self, jitter=0.0, seed=0):
        """Box building standing on the floor (its base sinks one cell into the floor)."""
        rng = np.random.default_rng(seed)
        j = 1.0 + jitter * (2.0 * rng.random(3) - 1.0) if jitter > 0 else np.ones(3)
        bx, bz, bh = self.bx * j[0], self.bz * j[1], self.bh * j[2]
        x0 = self.x_front * self.nx - 0.5 * self.nx
        floor = -0.5 * self.height
        return L.mesh_box((x0, floor - 1.0, -0.5 * bz), (x0 + bx, floor + bh, 0.5 * bz)), (bx, bz, bh)

    def context(self, storage=L.AERO_STORE_FP32, jitter=0.0, seed=0, closure=None):
        tris, dims = self.mesh(jitter, seed)
        ctx = L.Context(self.nx, self.height, self.lateral, storage=storage)
        ctx.set_mesh(tris, L.Tunnel(float(self.nx), float(self.lateral), float(self.height)))
        if closure is not None:
            ctx.set_closure(closure)
        ctx.set_params(L.SolverParams(u_inlet=self.u_ref, nu=self.nu, smagorinsky_cs=self.cs, ground_noslip=True,
                                      inflow_profile=L.AERO_PROFILE_LOGLAW, z0=self.z0, z_ref=self.z_ref,
                                      closure=L.AERO_CLOSURE_MLP if closure is not None else L.AERO_CLOSURE_SMAGORINSKY))
        ctx.reset(seed, 1e-4)
        st = L.mesh_stats(tris)
        refs = dict(frontal_area=dims[1] * dims[2], char_length=dims[0], tunnel_width=float(self.lateral),
                    tunnel_height=float(self.height), dx=1.0, u_ref=self.u_ref)
        return ctx, refs, st


def generate_samples(closure, case, sizes=4, snapshots_per_size=50, spinup=None, every=None, stride=2, jitter=0.1,
                     seed=100, log=print):
    """Runs the 2x-resolution plain (Smagorinsky) case for `sizes` seeded building sizes and appends, from every
    snapshot, samples filtered onto the coarse grid. Returns the number of samples added."""
    fine = case.scaled(2)
    spinup = spinup if spinup is not None else 2 * fine.nx
    every = every if every is not None else 20
    total = 0
    for s in range(sizes):
        ctx, _, _ = fine.context(jitter=jitter, seed=seed + s)
        ctx.step(spinup)
        for k in range(snapshots_per_size):
            ctx.step(every, readout=True)
            total += closure.sample_fine(ctx, stride, seed * 7919 + s * 131 + k, case.nu, case.cs)
        ctx.close()
        log(f"fine run {s + 1}/{sizes}: {total} samples")
    return total


def shuffle_samples(closure, seed=0):
    X, T = closure.samples()
    p = np.random.default_rng(seed).permutation(len(T))
    closure.clear_samples()
    closure.add_samples(X[p], T[p])
    return X[p], T[p]


def run_coefficients(case, steps, window, closure=None, seed=0, jitter=0.0, sample_every=1):
    """Runs a case and returns mean Cd and Cl (the uplift coefficient) over the last `window` steps."""
    ctx, refs, _ = case.context(jitter=jitter, seed=seed, closure=closure)
    ctx.step(steps - window)
    cd, cl = [], []
    for k in range(0, window, sample_every):
        ctx.step(sample_every)
        f = ctx.read_forces(refs)
        cd.append(f["cd"])
        cl.append(f["cl"])
    ctx.close()
    return float(np.mean(cd)), float(np.mean(cl))


def judge(case, closure, coarse_steps, window, seed=999, jitter=0.0):
    """Cd/Cl of the fine plain run (the judge) and of the coarse Smagorinsky and coarse learned runs."""
    fine = case.scaled(2)
    cd_f, cl_f = run_coefficients(fine, 2 * coarse_steps, 2 * window, seed=seed, jitter=jitter, sample_every=2)
    cd_s, cl_s = run_coefficients(case, coarse_steps, window, seed=seed, jitter=jitter)
    cd_l, cl_l = run_coefficients(case, coarse_steps, window, closure=closure, seed=seed, jitter=jitter)
    return {"fine": {"cd": cd_f, "cl": cl_f}, "smagorinsky": {"cd": cd_s, "cl": cl_s}, "learned": {"cd": cd_l, "cl": cl_l},
            "err_cd_smagorinsky": abs(cd_s - cd_f) / abs(cd_f), "err_cd_learned": abs(cd_l - cd_f) / abs(cd_f),
            "err_cl_smagorinsky": abs(cl_s - cl_f) / max(1e-9, abs(cl_f)),
            "err_cl_learned": abs(cl_l - cl_f) / max(1e-9, abs(cl_f))}
