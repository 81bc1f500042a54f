"""The Node binding loads without a GPU, exports GfProc, and turns bad
arguments and a missing device into JavaScript exceptions (the constructor
checks of src/gf.cc:253-400)."""
import json
import os
import shutil
import subprocess

import pytest

from parpar_amd import build as BLD

SCRIPT = r"""
const b = require(process.argv[1]);
const out = {type: typeof b.GfProc, errors: []};
const cases = [[7], [4096, {}], [4096, null, [{}]], [4096, null, null, 0],
               [4096, null, null, 2, [{slice_offset: 2, slice_size: 10}]],
               [4096, null, null, 2, [{slice_offset: 0, slice_size: 1001}, {slice_offset: 1001, slice_size: 3095}]]];
for (const args of cases) {
  try { new b.GfProc(...args); out.errors.push(null); } catch (e) { out.errors.push(e.message); }
}
console.log(JSON.stringify(out));
"""


def test_node_binding_loads_and_checks_arguments():
    if shutil.which("node") is None or not os.path.exists(BLD.NODE_ADDON):
        pytest.skip("node or the binding is not available")
    r = subprocess.run(["node", "-e", SCRIPT, BLD.NODE_ADDON], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["type"] == "function"
    e = out["errors"]
    assert e[0] == "Slice size is invalid"
    assert "CPU" in e[1] and "OpenCL" in e[2]
    assert e[3] == "Staging area count is invalid"
    assert e[4] == e[5] == "Slice portions allocated to the GPUs are invalid"  # gap; odd split not last


PLAN_SCRIPT = r"""
const P = require(process.argv[1]);
const b = require(process.argv[2]);
const GiB = 1 << 30;
const mems = {0: 288 * GiB, 1: 96 * GiB, 2: 288 * GiB};
const info = (d) => (d in mems ? {memory_global: mems[d]} : null);
const cases = [
  {sliceSize: 1 << 20, hipDevices: [{device: 0}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 0}, {device: 1}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 0, ratio: 0.25}, {device: 1}]},
  {sliceSize: 100000, hipDevices: [{device: 0, memoryLimit: 100 * GiB}, {device: 1, memoryLimit: GiB}]},
  {sliceSize: 40000, hipDevices: [{device: 0}, {device: 1}]},
  {sliceSize: 1000002, hipDevices: [{device: 0}, {device: 1}, {device: 2}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 0, ratio: 0.25}, {device: 1, ratio: 0.5}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 0, ratio: 0.6}, {device: 1, ratio: 0.5}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 0}, {device: 0}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 5}]},
  {sliceSize: 1 << 20, hipDevices: [{device: 0, minChunkSize: 3}]},
  {sliceSize: 5000, hipDevices: [{device: 0, input_batchsize: 2}]},
];
const out = cases.map((o) => { try { return P.planHip(o, info); } catch (e) { return e.message; } });
let noDevice;
try { b.hip_device_info(0); noDevice = null; } catch (e) { noDevice = e.message; }
console.log(JSON.stringify({plans: out, devices: b.hip_devices(), noDevice: noDevice}));
"""


def test_planner_divides_slices_like_the_opencl_planner():
    """Shares from ratios, else from memory limits (80 % of device memory by
    default), raised to the minimum chunk; devices dropped when the minimum
    chunks overflow the slice; even offsets; no CPU share, so the ranges cover
    the whole slice (lib/par2gen.js:488-607,680-694).  Expected ranges are
    worked out by hand from those rules."""
    if shutil.which("node") is None or not os.path.exists(BLD.NODE_ADDON):
        pytest.skip("node or the binding is not available")
    plan_js = os.path.join(os.path.dirname(BLD.NODE_ADDON), "..", "node", "plan.js")
    r = subprocess.run(["node", "-e", PLAN_SCRIPT, plan_js, BLD.NODE_ADDON], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    p = out["plans"]
    ranges = lambda k: [(d["device"], d["slice_offset"], d["slice_size"]) for d in p[k]["devices"]]
    GiB = 1 << 30
    assert ranges(0) == [(0, 0, 1 << 20)]
    assert p[0]["minMemoryLimit"] == int(288 * GiB * 0.8)
    assert ranges(1) == [(0, 0, 786432), (1, 786432, 262144)]           # 288 : 96 GiB = 3 : 1
    assert ranges(2) == [(0, 0, 262144), (1, 262144, 786432)]           # the rest to the device without a ratio
    assert ranges(3) == [(0, 0, 67232), (1, 67232, 32768)]              # 1 % share raised to the 32 KiB minimum
    assert ranges(4) == [(0, 0, 40000)]                                 # two 32 KiB minimums exceed the slice
    # 288 : 96 : 288 GiB = 3 : 1 : 3 of 1000002, rounded to 2 bytes, the last taking the rest
    assert ranges(5) == [(0, 0, 428572), (1, 428572, 142858), (2, 571430, 428572)]
    assert "cover the whole slice" in p[6]
    assert "exceed 100%" in p[7]
    assert "more than once" in p[8]
    assert "device info" in p[9]
    assert "min chunk size" in p[10]
    assert p[11]["devices"] == [{"device": 0, "slice_offset": 0, "slice_size": 5000, "input_batchsize": 2}]
    if not out["devices"]:  # no GPU in this process
        assert out["noDevice"] == "Invalid HIP device"
