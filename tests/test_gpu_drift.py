"""Precision form on the finned rocket: the shifted-FP16 population store against the plain FP32 run and the
FP64 oracle. At 256x128x128 the oracle's coefficients come from the committed fixture
tests/golden/rocket256_cd_fp64.json (tests/golden/make_rocket256_cd.py, about 10 CPU minutes); at the full
1024x512x512 size FP16 is compared with FP32 on the GPU. The bar is |dCd| <= 0.5 % on the mean Cd over the
last 400 of 2000 steps."""
import json
import os

import numpy as np
import pytest

import aerosim_amd as A
from bench import rocket_workload

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rocket256_cd_fp64.json")


def _run(nx, ny, nz, scale, storage, steps, window, u=0.08, re=2e4, cs=0.17, seed=7, amp=1e-4):
    tris, D = rocket_workload(nx, ny, nz, scale)
    ctx = A.Context(nx, ny, nz, storage=storage)
    ctx.set_mesh(tris, A.Tunnel(float(nx), float(nz), float(ny)))
    ctx.set_params(A.SolverParams(u_inlet=u, nu=u * D / re, smagorinsky_cs=cs))
    ctx.reset(seed, amp)
    st = A.mesh_stats(tris)
    refs = dict(frontal_area=st["frontal_area"], char_length=st["characteristic_length"], tunnel_width=float(nz),
                tunnel_height=float(ny), dx=1.0, u_ref=u)
    ctx.step(steps - window)
    rows = []
    for _ in range(window):
        ctx.step(1)
        f = ctx.read_forces(refs)
        rows.append([f["cd"], f["cl"], f["cm"]])
    ctx.close()
    return np.array(rows).mean(axis=0)


@pytest.mark.skipif(not os.path.exists(FIXTURE), reason="FP64 fixture not generated (tests/golden/make_rocket256_cd.py)")
def test_rocket256_fp32_and_fp16_against_fp64_oracle():
    fx = json.load(open(FIXTURE))
    c = fx["case"]
    ref = fx["mean"]
    res = {}
    for name, st in (("fp32", A.AERO_STORE_FP32), ("fp16", A.AERO_STORE_FP16)):
        res[name] = _run(c["nx"], c["ny"], c["nz"], c["scale"], st, c["steps"], c["window"])
    d = {k: abs(v[0] - ref["cd"]) / abs(ref["cd"]) for k, v in res.items()}
    print(f"rocket 256x128x128 mean Cd: fp64 oracle {ref['cd']:.6f}, fp32 {res['fp32'][0]:.6f} "
          f"(|dCd| {d['fp32'] * 100:.3f} %), fp16 {res['fp16'][0]:.6f} (|dCd| {d['fp16'] * 100:.3f} %); "
          f"Cl {ref['cl']:.5f} / {res['fp32'][1]:.5f} / {res['fp16'][1]:.5f}; Cm {ref['cm']:.5f} / {res['fp32'][2]:.5f} / "
          f"{res['fp16'][2]:.5f}")
    assert d["fp32"] <= 0.005
    assert d["fp16"] <= 0.005


def test_rocket1024_fp16_against_fp32():
    res = {}
    for name, st in (("fp32", A.AERO_STORE_FP32), ("fp16", A.AERO_STORE_FP16)):
        res[name] = _run(1024, 512, 512, 2.0, st, 2000, 400)
    d = abs(res["fp16"][0] - res["fp32"][0]) / abs(res["fp32"][0])
    print(f"rocket 1024x512x512 mean Cd: fp32 {res['fp32'][0]:.6f}, fp16 {res['fp16'][0]:.6f}: |dCd| {d * 100:.3f} %; "
          f"Cl {res['fp32'][1]:.5f} / {res['fp16'][1]:.5f}")
    assert d <= 0.005
