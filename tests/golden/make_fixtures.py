"""Writes the regression fixtures of tests/golden from the CPU oracle.

The reference implementation (TypeScript, three.js) cannot run in this environment and holds no golden
vectors of its own, so these fixtures are the oracle's outputs on a small seeded case: they freeze the
restatement (tests/test_golden.py checks the oracle and the GPU against them) rather than pin it to the
reference. Case: 16x8x8 cells, level-2 icosphere of diameter 5 cells centred at x = 6, u = 0.06, nu = 0.003,
Cs = 0.17, seeded 1e-4 noise (seed 17); flags, q, populations after 1 and 10 steps in fp64 and ref32 modes,
forces after 10 steps.

Run: python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

CASE = dict(nx=16, ny=8, nz=8, level=2, radius=2.5, xc=6.0, u=0.06, nu=0.003, cs=0.17, seed=17, amp=1e-4)


def build_case():
    import aerosim_amd as A
    from oracle import oracle as O

    c = CASE
    tris = A.mesh_icosphere(c["level"], c["radius"], (c["xc"] - c["nx"] / 2, 0.0, 0.0))
    vox = O.Voxels(tris, float(c["ny"]), (c["nx"], c["ny"], c["nz"]))
    return tris, vox


def main():
    from oracle import oracle as O

    c = CASE
    tris, vox = build_case()
    out = dict(tris=tris, flags=vox.cell_flags, bmap=vox.boundary_index_map, blist=vox.boundary_cell_indices,
               q=vox.boundary_q_values)
    for mode in ("fp64", "ref32"):
        s = O.Solver(vox, mode, u_inlet=c["u"], nu=c["nu"], cs=c["cs"])
        s.reset(c["seed"], c["amp"])
        s.step(1)
        out[f"f1_{mode}"] = s.populations()
        s.step(9)
        out[f"f10_{mode}"] = s.populations()
        f = s.forces()
        out[f"forces10_{mode}"] = np.array([f[k] for k in ("fx", "fy", "fz", "mx", "my", "mz")])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere16.npz"), **out)
    print("wrote sphere16.npz")


if __name__ == "__main__":
    main()
