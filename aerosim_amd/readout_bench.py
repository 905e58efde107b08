"""Timing of the derived-field readouts on the finned-rocket tunnel (default 512x256x256, u = 0.08, Re 2e4,
Smagorinsky Cs = 0.17, seeded): after `steps` steps with a readout on the last one, each readout runs
`repeat` times and its mean wall time per call is reported. The Q-criterion is timed without the host copy
of its fields (maxima only) and with it; the slice is timed on all three axes.

Run: python -m aerosim_amd.readout_bench [--grid 512,256,256] [--steps 50] [--repeat 20]
Prints one JSON object. For per-kernel durations, run it under `rocprofv3 --kernel-trace --stats`.
"""
import argparse
import json
import time

import numpy as np

from . import lbm as L

# Algorithmic HBM bytes per cell of the Q-criterion kernel: ux, uy, uz and the flag read once (13 B),
# Q, |omega| and the 3-float vorticity written (20 B).
Q_BYTES_PER_CELL = 33


def _rocket(nx, ny, nz, scale, seed=7):
    D, Ln, Lb = 24.0 * scale, 60.0 * scale, 160.0 * scale
    tris = L.mesh_rocket((100.0 * scale - 0.5 * nx, 0.0, 0.0), Lb, D, Ln, 0.02, seed, 128)
    return tris, D


def _time(fn, repeat):
    fn()
    t0 = time.perf_counter()
    for _ in range(repeat):
        fn()
    return (time.perf_counter() - t0) / repeat * 1e3


def run(grid=(512, 256, 256), steps=50, repeat=20, log=None):
    nx, ny, nz = grid
    tris, D = _rocket(nx, ny, nz, nx / 512.0)
    ctx = L.Context(nx, ny, nz)
    ctx.set_mesh(tris, L.Tunnel(float(nx), float(nz), float(ny)))
    u = 0.08
    ctx.set_params(L.SolverParams(u_inlet=u, nu=u * D / 2e4, smagorinsky_cs=0.17))
    ctx.reset(7, 1e-4)
    ctx.step(steps, readout=True)
    ctx.sync()
    if log:
        log(f"stepped {steps} on {nx}x{ny}x{nz}")
    N = nx * ny * nz
    res = {"grid": [nx, ny, nz], "cells": N, "steps": steps, "repeat": repeat}
    q_ms = _time(lambda: ctx.compute_q_criterion(1.0, fields=False), repeat)
    res["qcriterion_ms"] = q_ms
    res["qcriterion_gbps_wall"] = N * Q_BYTES_PER_CELL / (q_ms * 1e-3) / 1e9
    res["qcriterion_with_copy_ms"] = _time(lambda: ctx.compute_q_criterion(1.0, fields=True), max(2, repeat // 5))
    r = ctx.compute_q_criterion(1.0, fields=False)
    res["max_q"], res["max_omega"] = r["max_q"], r["max_omega"]
    for axis in ("z", "y", "x"):
        res[f"slice_{axis}_ms"] = _time(lambda: ctx.extract_slice_rgba(axis, 0.5, "vorticity", "cpu", u), repeat)
    pts = np.random.default_rng(1).uniform(-0.5, 0.5, (4096, 3)) * np.array([nx, ny, nz])
    res["probe_4096_ms"] = _time(lambda: ctx.probe(pts, (-0.5 * nx, -0.5 * ny, -0.5 * nz), 1.0, 25.0, u), repeat)
    ctx.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", default="512,256,256")
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--repeat", type=int, default=20)
    a = ap.parse_args()
    grid = tuple(int(v) for v in a.grid.split(","))
    print(json.dumps(run(grid, a.steps, a.repeat, log=lambda m: print(m, flush=True))))


if __name__ == "__main__":
    main()
