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
              "bakeSdf"):
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
