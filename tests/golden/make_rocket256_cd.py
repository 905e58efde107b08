"""Writes tests/golden/rocket256_cd_fp64.json: the FP64 oracle's force coefficients for the finned-rocket
case at 256x128x128 cells (the 1024x512x512 precision-form workload at a quarter of its linear size), the
reference against which the GPU's FP32 and shifted-FP16 runs are compared (tests/test_gpu_drift.py).
With --size 128 it writes tests/golden/rocket128_cd_fp64.json instead: the 512x256x256 plain workload at a
quarter of its linear size (rocket scaled by 1/4: D = 6, L = 40 cells), about 1 CPU minute on 8 cores.

Case: rocket of the bench workload scaled by 1/2 (D = 12, L = 80 cells, ogive nose 30 cells), u = 0.08,
Re = 2e4 (nu = u D / Re), Smagorinsky Cs = 0.17, seeded 1e-4 noise (seed 7); 2000 steps; Cd, Cl, Cm and the
centre of pressure at every step of the last 400. Takes about 10 minutes of CPU time on 8 cores.

Run: python tests/golden/make_rocket256_cd.py [--size 128]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

CASES = {
    256: dict(nx=256, ny=128, nz=128, scale=0.5, u=0.08, re=2e4, cs=0.17, seed=7, amp=1e-4, steps=2000, window=400),
    128: dict(nx=128, ny=64, nz=64, scale=0.25, u=0.08, re=2e4, cs=0.17, seed=7, amp=1e-4, steps=2000, window=400),
}


def main():
    import aerosim_amd as A
    from bench import rocket_workload
    from oracle import oracle as O

    size = int(sys.argv[sys.argv.index("--size") + 1]) if "--size" in sys.argv else 256
    c = CASES[size]
    OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"rocket{size}_cd_fp64.json")
    tris, D = rocket_workload(c["nx"], c["ny"], c["nz"], c["scale"])
    vox = O.Voxels(tris, float(c["ny"]), (c["nx"], c["ny"], c["nz"]))
    s = O.Solver(vox, "fp64", u_inlet=c["u"], nu=c["u"] * D / c["re"], cs=c["cs"])
    s.reset(c["seed"], c["amp"])
    st = A.mesh_stats(tris)
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=float(c["nz"]),
                tunnel_height=float(c["ny"]), dx=1.0, u_ref=c["u"])
    t0 = time.perf_counter()
    s.step(c["steps"] - c["window"])
    rows = []
    for _ in range(c["window"]):
        s.step(1)
        k = O.coefficients(s.forces(), refs["frontal_area"], refs["char_length"], refs["tunnel_width"],
                           refs["tunnel_height"], refs["dx"], refs["u_ref"])
        rows.append([k["cd"], k["cl"], k["cm"], k["xcp"]])
    a = np.array(rows)
    out = {"case": c, "refs": refs, "triangles": int(tris.shape[0]), "cd": a[:, 0].tolist(), "cl": a[:, 1].tolist(),
           "cm": a[:, 2].tolist(), "xcp": a[:, 3].tolist(), "mean": {"cd": float(a[:, 0].mean()), "cl": float(a[:, 1].mean()),
                                                                   "cm": float(a[:, 2].mean()), "xcp": float(a[:, 3].mean())},
           "oracle_seconds": time.perf_counter() - t0}
    with open(OUT, "w") as fh:
        json.dump(out, fh)
    print("wrote", OUT, out["mean"])


if __name__ == "__main__":
    main()
