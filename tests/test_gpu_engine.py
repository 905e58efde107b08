"""The engine facades on the GPU: the reference's frame loop (main.ts:491-512: 4 substeps per frame, MLUPS
from the frame time) through the Python and Node LBMEngine, and the CSV / JSON reports of the export
buttons (main.ts:294-385) from the engine's latest force readout."""
import json
import os
import shutil
import subprocess

import pytest

import aerosim_amd as A

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _engine():
    tunnel = A.Tunnel(3.0, 1.5, 1.5)
    vox = A.voxelize_mesh_to_lattice(A.mesh_icosphere(2, 0.3, (0, 0, 0)), None, tunnel, (48, 24, 24))
    eng = A.LBMEngine()
    eng.update_config(u_inlet_lattice=0.07, nu_lattice=0.004, smagorinsky_cs=0.12, tunnel=tunnel, wind_speed_mps=30.0,
                      frontal_area_m2=0.2827, characteristic_length_m=0.6)
    eng.set_voxel_grid(vox)
    return eng


def test_python_engine_frames_and_reports():
    eng = _engine()
    frames = [eng.tick() for _ in range(30)]
    assert [f["stepCount"] for f in frames] == [4 * (i + 1) for i in range(30)]
    assert all(f["mlups"] > 0 and f["width"] == 48 and f["height"] == 24 for f in frames)
    last = frames[-1]["forceReadout"]
    assert len(last["history"]) == 30 and last["history"][-1]["step"] == 120
    print(f"frame MLUPS (48x24x24, 4 substeps, readout every frame): {frames[-1]['mlups']:.0f}; "
          f"mean Cd {last['meanCd']:.4f}")
    csv = eng.export_telemetry_csv("Sphere", timestamp="2026-01-01T00:00:00.000Z").split("\r\n")
    assert csv[1] == "# Generated: 2026-01-01T00:00:00.000Z" and len(csv) == 12 + 30
    assert csv[-1].startswith("120,0.1200,")
    rep = json.loads(eng.export_run_report_json("Sphere", timestamp="2026-01-01T00:00:00.000Z"))
    run = rep["simulationRun"]
    assert run["totalSteps"] == 120 and len(run["telemetryHistory"]) == 30 and run["specimenName"] == "Sphere"
    assert abs(run["reynolds"] - 30.0 * 0.6 / 1.516e-5) < 1e-6 * run["reynolds"]
    eng.ctx.close()


def test_node_engine_frames():
    node = shutil.which("node")
    if node is None or not os.path.exists(os.path.join(ROOT, "aerosim_amd", "node", "aerolbm.node")):
        pytest.skip("node or the addon is absent")
    script = """
const m = require('./aerosim_amd/node/lbmEngine.js');
const tunnel = {length: 3.0, width: 1.5, height: 1.5};
const vox = m.voxelizeMeshToLattice({position: m.addon.meshIcosphere(2, 0.3, [0, 0, 0])}, null, tunnel, {nx: 48, ny: 24, nz: 24});
const e = new m.LBMEngine(0);
e.setVoxelGrid(vox, {uInletLattice: 0.07, nuLattice: 0.004, smagorinskyCs: 0.12, tunnel, windSpeedMps: 30.0});
(async () => {
  const out = [];
  for (let i = 0; i < 10; i++) out.push(await e.tick(i === 9));
  const last = out[9];
  console.log(JSON.stringify({steps: out.map((f) => f.stepCount), mlups: out.map((f) => f.mlups),
    hist: last.forceReadout.history.length, macro: last.macroField ? last.macroField.ux.length : 0,
    csvRows: m.exportTelemetryToCSV(last.forceReadout.history.map((h) => ({step: h.step, timeSec: h.step * 0.001,
      cd: h.cd, cl: h.cl, cm: h.cm, dragForceN: h.dragForceN, liftForceN: h.liftForceN, xcpMeters: h.xcp})),
      {timestamp: 'T'}).split('\\r\\n').length}));
  e.destroy();
})();
"""
    p = subprocess.run([node, "-e", script], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    r = json.loads(p.stdout)
    assert r["steps"] == [4 * (i + 1) for i in range(10)] and all(v > 0 for v in r["mlups"])
    assert r["hist"] == 10 and r["macro"] == 48 * 24 * 24 and r["csvRows"] == 22


def test_node_bindings_size_arrays_from_the_context():
    node = shutil.which("node")
    if node is None or not os.path.exists(os.path.join(ROOT, "aerosim_amd", "node", "aerolbm.node")):
        pytest.skip("node or the addon is absent")
    script = """
const a = require('./aerosim_amd/node/aerolbm.node');
const ctx = a.createContext(16, 8, 8, 0, 0);
const n = 16 * 8 * 8;
const errs = [];
function expectType(f) { try { f(); errs.push('no error'); } catch (e) { errs.push(e instanceof TypeError ? 'TypeError' : String(e)); } }
expectType(() => a.setVoxels(ctx, new Uint8Array(n - 1), new Int32Array(n), new Int32Array(0), new Float32Array(0)));
expectType(() => a.setVoxels(ctx, new Uint8Array(n), new Int32Array(n - 1), new Int32Array(0), new Float32Array(0)));
expectType(() => a.setVoxels(ctx, new Uint8Array(n), new Int32Array(n).fill(-1), new Int32Array(0), new Float32Array(18)));
expectType(() => a.readMacro(ctx, n - 1));
expectType(() => a.readPopulations(ctx, 1));
expectType(() => a.readVoxelArrays(ctx, 2 * n));
a.setVoxels(ctx, new Uint8Array(n), new Int32Array(n).fill(-1), new Int32Array(0), new Float32Array(0));
console.log(JSON.stringify(errs));
a.destroy(ctx);
"""
    p = subprocess.run([node, "-e", script], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert json.loads(p.stdout) == ["TypeError"] * 6
