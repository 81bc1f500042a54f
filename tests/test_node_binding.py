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
