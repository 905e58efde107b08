"""The Node host (aerosim_amd/node): the N-API addon loads and exposes the bindings; without a GPU the engine
constructor path throws instead of falling back to a CPU solver."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "aerosim_amd", "node", "aerolbm.node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon is absent")


def _node(script):
    p = subprocess.run([NODE, "-e", script], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p.stdout


def test_addon_exports_and_host_geometry():
    out = _node("""
const m = require('./aerosim_amd/node/lbmEngine.js');
const t = m.addon.meshIcosphere(3, 4.0, [0, 0, 0]);
console.log(JSON.stringify({keys: Object.keys(m.addon), ntri: t.length / 9, area: m.addon.meshStats(t).frontalArea}));
""")
    r = json.loads(out)
    for k in ("createContext", "setMesh", "setVoxels", "setParams", "reset", "step", "readForces", "readMacro",
              "readPopulations", "readVoxelArrays", "destroy"):
        assert k in r["keys"]
    assert r["ntri"] == 1280
    assert 49.0 < r["area"] < 50.3


def test_engine_without_gpu_throws():
    import aerosim_amd as A

    try:
        A.Context(8, 8, 8).close()
        pytest.skip("a GPU is present")
    except A.AeroError:
        pass
    out = _node("""
const m = require('./aerosim_amd/node/lbmEngine.js');
try { m.voxelizeMeshToLattice({position: m.addon.meshIcosphere(1, 2.0, [0,0,0])}, null,
      {length: 16, width: 8, height: 8}, {nx: 16, ny: 8, nz: 8}); console.log('no-throw'); }
catch (e) { console.log('threw ' + e.message); }
""")
    assert out.startswith("threw") and "aero_create" in out


def test_stl_through_the_node_host():
    """parseSTL / normalizeAndCenterGeometry from JavaScript give the bits of the Python binding."""
    import numpy as np

    import aerosim_amd as A
    from tests.test_stl import TETRA

    out = _node("""
const m = require('./aerosim_amd/node/lbmEngine.js');
const text = %s;
const g = m.parseSTL(Buffer.from(text, 'utf8'));
const bin = Buffer.alloc(84 + 50);
bin.writeUInt32LE(1, 80);
[0, 0, 0, 120, 0, 0, 0, 60, 30].forEach((v, i) => bin.writeFloatLE(v, 84 + 12 + 4 * i));
const b = m.parseSTL(bin.buffer.slice(bin.byteOffset, bin.byteOffset + bin.length));
const n = m.normalizeAndCenterGeometry(b);
let err = '';
try { m.parseSTL(new Uint8Array(100)); } catch (e) { err = e.message; }
console.log(JSON.stringify({tetra: Array.from(g.position), norm: Array.from(n.geometry.position),
                            scaled: n.autoScaledFromMm, err}));
""" % json.dumps(TETRA))
    r = json.loads(out)
    assert np.array_equal(np.array(r["tetra"], dtype=np.float32), A.parse_stl(TETRA.encode()).reshape(-1))
    want, scaled = A.normalize_and_center_geometry(np.array([[0, 0, 0, 120, 0, 0, 0, 60, 30]], np.float32))
    assert r["scaled"] == scaled and np.array_equal(np.array(r["norm"], dtype=np.float32), want.reshape(-1))
    assert "Invalid or corrupted STL file." in r["err"]
