"""Grid sizing and force/coefficient helpers of the reference's host (src/core/gpuContext.ts:10-58,
src/solver/fluidProperties.ts:7-80), with the reference's own cases (src/validation/phase1.test.ts:100-136).
CPU only."""
import json
import math
import os
import shutil
import subprocess

import pytest

from aerosim_amd import export as E
from aerosim_amd import sizing as S

AIR = {"density": 1.204, "kinematicViscosity": 1.516e-5, "speedOfSound": 343.2}


def test_phase1_telemetry_and_coefficient_round_trip():
    """phase1.test.ts:100-121: Re = 30 / 1.516e-5, q = 541.8 Pa, supercritical regime; Cd 0.47 back from its
    force to 5 decimals."""
    t = E.compute_aerodynamic_telemetry(30.0, 1.0, 0.7854, AIR, {"length": 6, "width": 2.5, "height": 2.5})
    assert abs(t["reynoldsNumber"] - 30 / 1.516e-5) < 0.5
    assert abs(t["dynamicPressurePa"] - 541.8) < 0.005
    assert t["flowRegime"] == "Turbulent (Supercritical)"
    f = 0.47 * t["dynamicPressurePa"] * 0.7854
    assert abs(S.force_to_coefficient(f, AIR["density"], 30.0, 0.7854) - 0.47) < 5e-6
    assert abs(S.coefficient_to_force(0.47, AIR["density"], 30.0, 0.7854) - f) < 1e-9
    assert S.force_to_coefficient(1.0, 1.204, 0.0, 1.0) == 0.0


def test_phase1_safe_grid_resolution():
    """phase1.test.ts:123-135: 8-aligned sizes, nx > ny, totalCells their product. For these inputs the cell
    pitch is exactly 0.0375 m (160 x 80 x 80 = the budget), so the grid is 160 x 80 x 80 and the reference's
    estimate 1 024 000 x 216 B = 210.9 MiB."""
    g = S.calculate_safe_grid_resolution(128 * 1024 * 1024, 256 * 1024 * 1024, {"length": 6, "width": 3, "height": 3},
                                         160 * 80 * 80)
    assert g["nx"] % 8 == 0 and g["ny"] % 8 == 0 and g["nz"] % 8 == 0
    assert g["nx"] > g["ny"] and g["totalCells"] == g["nx"] * g["ny"] * g["nz"]
    assert (g["nx"], g["ny"], g["nz"], g["estimatedVramMB"]) == (160, 80, 80, 210.9)


def test_safe_grid_resolution_shrinks_to_the_buffer_limit():
    """A budget far beyond a 64 MiB buffer: the single population buffer ends under 90 % of the limit, with
    the minimum sizes kept."""
    lim = 64 * 1024 * 1024
    g = S.calculate_safe_grid_resolution(lim, 2 * lim, {"length": 8, "width": 2, "height": 2}, 10 ** 9)
    assert g["totalCells"] * 76 <= 0.9 * lim
    assert g["nx"] >= 64 and g["ny"] >= 32 and g["nz"] >= 32
    cap, _ = S.mi355x_buffer_limits()
    big = S.calculate_safe_grid_resolution(cap, cap, {"length": 6, "width": 3, "height": 3}, 512 * 256 * 256)
    assert big["totalCells"] <= 512 * 256 * 256 and big["totalCells"] < 2 ** 31


@pytest.mark.skipif(shutil.which("node") is None, reason="node is absent")
def test_node_host_helpers_match_python():
    cases = [[128 * 2 ** 20, 256 * 2 ** 20, {"length": 6, "width": 3, "height": 3}, 1024000],
             [64 * 2 ** 20, 128 * 2 ** 20, {"length": 8, "width": 2, "height": 2}, 10 ** 9],
             [2 ** 30, 2 ** 31, {"length": 4.5, "width": 2.8, "height": 2.4}, 3_000_000],
             [2 ** 28, 2 ** 28, {"length": 0.05, "width": 7, "height": 1.3}, 50_000]]
    script = ("const m = require('./aerosim_amd/node/lbmEngine.js');"
              "const cases = JSON.parse(process.argv[1]);"
              "console.log(JSON.stringify({g: cases.map((c) => m.calculateSafeGridResolution(...c)),"
              " c: m.forceToCoefficient(215.5, 1.204, 30, 0.7854), f: m.coefficientToForce(0.47, 1.204, 30, 0.7854)}));")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run(["node", "-e", script, json.dumps(cases)], capture_output=True, text=True, timeout=60, cwd=root)
    assert p.returncode == 0, p.stderr
    out = json.loads(p.stdout)
    assert out["g"] == [S.calculate_safe_grid_resolution(*c) for c in cases]
    assert out["c"] == S.force_to_coefficient(215.5, 1.204, 30, 0.7854)
    assert out["f"] == S.coefficient_to_force(0.47, 1.204, 30, 0.7854)
    assert math.isfinite(out["c"])
