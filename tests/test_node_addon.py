"""The N-API binding (f1-wind-tunnel_amd/node) on the CPU: the addon loads and exports
its functions, argument errors become thrown Errors, and the JavaScript packing of the
Shape uniform is byte-identical to the Python host's (both follow engine.ts:429-448)."""
import ctypes as C
import json
import shutil
import subprocess
from pathlib import Path

import pytest

import aerolbm

ROOT = Path(__file__).resolve().parent.parent
NODE_DIR = ROOT / "f1-wind-tunnel_amd" / "node"
ADDON = ROOT / "f1-wind-tunnel_amd" / "lib" / "aero_napi.node"

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not ADDON.exists(),
                                reason="node or the built addon is absent")


def node(script: str) -> str:
    r = subprocess.run(["node", "-e", script], cwd=NODE_DIR, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_addon_exports_and_errors():
    out = json.loads(node("""
const { addon, HipAeroEngine } = require('./hip-engine');
let msg = '';
try { new HipAeroEngine({dims: [16, 8, 8], h: 0.1}, {tau0: 0.5}); } catch (e) { msg = e.message; }
let typeMsg = '';
try { addon.step({}, 1); } catch (e) { typeMsg = e.message; }
console.log(JSON.stringify({ keys: Object.keys(addon), version: addon.version(), msg, typeMsg }));
"""))
    for k in ("create", "destroy", "init", "voxelize", "setFlags", "step", "heatStep", "readForces", "readFields",
              "readTemps", "readFlags", "readPopulations", "guardCount", "sync", "writeFields", "sampleFlow",
              "particlesAlloc", "particlesStep", "readParticles", "writeParticles", "streamlines", "batchCreate",
              "batchInstances", "batchDestroy", "batchStep", "batchRunGraph", "batchAutotune", "batchSetLaunchShape",
              "bakeSdf", "vlmLoads"):
        assert k in out["keys"]
    assert out["version"] == aerolbm.load().aero_version()
    assert "aero_create failed" in out["msg"] and "tau0" in out["msg"]
    assert "context" in out["typeMsg"]


@pytest.mark.parametrize("params", [
    dict(kind=0, frontFlapDeg=2, rearFlapDeg=-3, rideH=0.11, rearRideH=0.13, aoaDeg=5, cooling=2,
         heatBrakes=True, heatEngine=False, heatRad=True, frontWing=1, rearWing=2, sidepod=3, cover=1, nose=2),
    dict(kind=4, customMode=1, customWidth=0.4, customAxisY=0.9, rideH=0.1, rearRideH=0.12, cooling=0,
         heatBrakes=False, heatEngine=True, heatRad=False),
])
def test_shape_packing_matches_python_host(params):
    js = node(f"""
const {{ packShape }} = require('./hip-engine');
const b = packShape({json.dumps(params)});
console.log(Buffer.from(b).toString('hex'));
""").strip()
    p = aerolbm.ShapeParams(**{k: v for k, v in params.items()})
    py = bytes(aerolbm.shape_uniform(p))
    assert bytes.fromhex(js) == py


@pytest.mark.parametrize("kind,kmh,mode", [(0, 250.0, 2), (2, 180.0, 3), (5, 320.0, 1), (1, 90.0, 4), (5, 250.0, 4),
                                           (0, 300.0, 4)])
def test_flow_uniforms_match_python_host(kind, kmh, mode):
    """viewUniform, the frame's PSim and streamConfig packed by the JavaScript host are
    byte-identical to the Python host's (engine.ts:503-572,595-641,731-737)."""
    from aerolbm import flow as FL
    p = aerolbm.ShapeParams(kind=kind, rideH=0.11, rearRideH=0.15)
    h = 0.045
    js = json.loads(node(f"""
const m = require('./hip-engine');
const self = {{ shapeParams: Object.assign({{}}, {{kind: {kind}, rideH: 0.11, rearRideH: 0.15, heatBrakes: true,
               heatEngine: true, heatRad: true, customWidth: 0.6, frontWingOn: true, rearWingOn: true}}), h: {h} }};
self.emitterPoints = m.HipAeroEngine.prototype.emitterPoints;
const sp = m.HipAeroEngine.prototype.spawnRegion.call(self);
const view = m.viewUniform(0.0925, {kmh}, self.shapeParams);
const rate = Math.min(1.6, Math.max(0.35, {kmh} / 250));
const ps = m.packPSim({{count: 140000, spawnMode: {mode} === 3 ? 0 : 1, colorMode: 0, frame: 7, dt: 8 * rate, life: 900,
                       scatter: true, rakeRows: sp.rows, rakeCols: sp.cols, spawnMin: sp.min, spawnMax: sp.max}});
const sc = m.HipAeroEngine.prototype.streamConfig.call(self, {mode} === 4 ? 4 : 1);
const st = m.packStream(Object.assign({{}}, sc, {{steps: m.STREAM_STEPS, colorMode: 0, time: 0.5}}));
console.log(JSON.stringify({{view: Buffer.from(view.buffer).toString('hex'), psim: Buffer.from(ps).toString('hex'),
                             stream: Buffer.from(st).toString('hex'), em: Buffer.from(sc.e.buffer).toString('hex')}}));
"""))
    view = FL.view_uniform(0.0925, kmh, p)
    assert bytes.fromhex(js["view"]) == bytes(view)
    ps = FL.psim_uniform(140000, FL.FlowMode.Fog if mode == 3 else FL.FlowMode.Smoke, 0, 7, 8, h, p, kmh,
                         scatter=True)
    assert bytes.fromhex(js["psim"]) == bytes(ps)
    st, em = FL.stream_uniform(p, FL.FlowMode.Vortices if mode == 4 else FL.FlowMode.Streamlines, 0, 0.5)
    assert bytes.fromhex(js["stream"]) == bytes(st)
    assert bytes.fromhex(js["em"]) == em.tobytes()


W17_SETUPS = [
    dict(frontDeg=0, rearFlapDeg=0, rearMainDeg=0, rideH=0.12, rearRideH=0.14, gurney=False, frontOn=True,
         rearOn=True, rearTwistDeg=0, flapScale=1),
    dict(frontDeg=8.5, rearFlapDeg=-15, rearMainDeg=-3.5, rideH=0.08, rearRideH=0.26, gurney=True, frontOn=True,
         rearOn=True, rearTwistDeg=-16, flapScale=0.72),
    dict(frontDeg=-9, rearFlapDeg=9, rearMainDeg=3, rideH=0.2, rearRideH=0.1, gurney=False, frontOn=True,
         rearOn=True, rearTwistDeg=6, flapScale=1),
]


def test_w17_lattices_match_python_host(tmp_path):
    """The JavaScript lattice builder (V8's Math functions, as the reference runs) and
    the Python one (the C library's cos / sin / atan) give the same horseshoes: bit for
    bit but for rare last-bit differences of those functions."""
    import numpy as np

    from aerolbm import vlm as V

    rig_path = ROOT / "tests" / "golden" / "w17_rig.json"
    out = tmp_path / "lat.f64"
    node(f"""
const fs = require('fs');
const {{ w17Lattices }} = require('./hip-engine');
const rig = JSON.parse(fs.readFileSync({json.dumps(str(rig_path))}));
const parts = [];
for (const s of {json.dumps(W17_SETUPS)}) {{ const l = w17Lattices(rig, s); parts.push(l.front, l.rear); }}
const o = new Float64Array(parts.reduce((a, p) => a + p.length, 0));
let k = 0;
for (const p of parts) {{ o.set(p, k); k += p.length; }}
fs.writeFileSync({json.dumps(str(out))}, Buffer.from(o.buffer));
""")
    got = np.fromfile(out, np.float64)
    rig = json.loads(rig_path.read_text())
    ref = np.concatenate([np.concatenate(V.w17_lattices(rig, V.W17WingSetup(**s))).ravel() for s in W17_SETUPS])
    assert got.shape == ref.shape == (((240 + 240) * 3 + 80) * 20,)
    differ = got.view(np.uint64) != ref.view(np.uint64)
    assert differ.sum() <= 0.001 * got.size
    assert np.abs(got - ref).max() <= 4 * np.finfo(np.float64).eps * max(1.0, np.abs(ref).max())
