"""Angle-of-attack / specimen sweep on the batched launch: many uneven tunnels, each holding an extruded
symmetric 4-digit-formula airfoil, stepped together by aero_batch (one bulk launch and one fluid-list launch
per step over all tunnels, the Bouzidi list on a side stream, graph-captured), against the same tunnels stepped
one at a time.

Workload (defaults): 64 tunnels, sizes uniform between 96^3 and 256x128x128 cells (every side a multiple of
4), airfoil thickness t/c uniform in [0.08, 0.18], chord uniform in [40, 80] cells, angle of attack uniform
in [-4, 16] degrees, inflow u = 0.06, Re uniform in [5e3, 5e4], Smagorinsky Cs = 0.17, seeded.

Run: python -m aerosim_amd.sweeps [--tunnels 64] [--steps 200] [--storage fp32|fp16]
Prints one JSON object: aggregate MLUPS of the batch and of the one-at-a-time loop, the tuned launch shape,
and per-tunnel Cl/Cd of both runs.
"""
import argparse
import json
import time

import numpy as np

from . import lbm as L


def sweep_cases(n=64, seed=2024):
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        nx = 4 * int(rng.integers(24, 65))      # 96 .. 256
        ny = 4 * int(rng.integers(24, 33))      # 96 .. 128
        nz = 4 * int(rng.integers(24, 33))
        cases.append(dict(nx=nx, ny=ny, nz=nz, t=float(rng.uniform(0.08, 0.18)), chord=float(rng.uniform(40.0, 80.0)),
                          alpha=float(rng.uniform(-4.0, 16.0)), re=float(10 ** rng.uniform(np.log10(5e3), np.log10(5e4))),
                          u=0.06, seed=int(rng.integers(1, 1 << 30))))
    return cases


def make_tunnel(c, storage=L.AERO_STORE_FP32):
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    chord = min(c["chord"], 0.45 * nx)
    tris = L.mesh_naca(c["t"], chord, (-0.3 * nx, 0.0, 0.0), c["alpha"], -nz, nz, 64)
    ctx = L.Context(nx, ny, nz, storage=storage)
    ctx.set_mesh(tris, L.Tunnel(float(nx), float(nz), float(ny)))
    ctx.set_params(L.SolverParams(u_inlet=c["u"], nu=c["u"] * chord / c["re"], smagorinsky_cs=0.17))
    ctx.reset(c["seed"], 1e-4)
    # reference quantities of a spanning airfoil: chord x span for the area, chord for the length
    refs = dict(frontal_area=chord * nz, char_length=chord, tunnel_width=float(nz), tunnel_height=float(ny), dx=1.0,
                u_ref=c["u"])
    return ctx, refs


def run_sweep(n=64, steps=200, storage=L.AERO_STORE_FP32, tune_steps=4, seed=2024, log=None, single=True,
              shape=None):
    cases = sweep_cases(n, seed)
    cells = sum(c["nx"] * c["ny"] * c["nz"] for c in cases)
    # batched
    tun = [make_tunnel(c, storage) for c in cases]
    ctxs = [t[0] for t in tun]
    b = L.Batch(ctxs)
    if shape is None:
        shape = b.tune(tune_steps)
    else:
        b.set_shape(shape["block"], shape["cells_per_thread"], shape["grid"])
    if log:
        log(f"shape {shape}")
    for (ctx, _), c in zip(tun, cases):  # tuning advanced the flow: start again from the seeded state
        ctx.reset(c["seed"], 1e-4)
    b.step(2)
    t_batch = b.time_steps(steps)
    for ctx, _ in tun:
        ctx.sync()
    batch_coef = [ctx.read_forces(refs) for ctx, refs in tun]
    b.close()
    for ctx, _ in tun:
        ctx.close()
    if not single:
        return {"tunnels": n, "cells": cells, "steps": steps, "batch": {"ms_per_step": t_batch["step_ms"],
                "mlups": cells / (t_batch["step_ms"] * 1e-3) / 1e6, "shape": shape}}
    # one at a time, same steps, same cells per thread
    singles = []
    t0 = time.perf_counter()
    total_single_ms = 0.0
    for c in cases:
        ctx, refs = make_tunnel(c, storage)
        ctx.set_cells_per_thread(shape["cells_per_thread"])
        ctx.step(2)
        tm = ctx.time_steps(steps)
        total_single_ms += tm["total_ms"]
        singles.append(ctx.read_forces(refs))
        ctx.close()
    wall_single = time.perf_counter() - t0
    per = []
    for c, fb, fs in zip(cases, batch_coef, singles):
        per.append({"size": [c["nx"], c["ny"], c["nz"]], "alpha": round(c["alpha"], 3), "t_c": round(c["t"], 4),
                    "re": round(c["re"]), "cd_batch": fb["cd"], "cl_batch": fb["cl"], "cd_single": fs["cd"],
                    "cl_single": fs["cl"], "identical": fb["fx"] == fs["fx"] and fb["fy"] == fs["fy"]})
    return {
        "tunnels": n, "cells": cells, "steps": steps, "storage": "fp16" if storage == L.AERO_STORE_FP16 else "fp32",
        "batch": {"ms_per_step": t_batch["step_ms"], "mlups": cells / (t_batch["step_ms"] * 1e-3) / 1e6,
                  "shape": shape},
        "one_at_a_time": {"ms_per_step": total_single_ms / steps, "mlups": cells / (total_single_ms / steps * 1e-3) / 1e6,
                          "wall_s": wall_single},
        "all_identical": all(p["identical"] for p in per),
        "per_tunnel": per,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tunnels", type=int, default=64)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--storage", choices=["fp32", "fp16"], default="fp32")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--shape", default=None, help="block,cells_per_thread,grid instead of autotuning")
    ap.add_argument("--batch-only", action="store_true", help="skip the one-at-a-time comparison")
    args = ap.parse_args()
    st = L.AERO_STORE_FP16 if args.storage == "fp16" else L.AERO_STORE_FP32
    shape = None
    if args.shape:
        bl, cpt, gr = (int(v) for v in args.shape.split(","))
        shape = {"block": bl, "cells_per_thread": cpt, "grid": gr}
    r = run_sweep(args.tunnels, args.steps, st, seed=args.seed, log=lambda m: print(m, flush=True),
                  single=not args.batch_only, shape=shape)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
