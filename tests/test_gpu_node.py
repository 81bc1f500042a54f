"""The Node binding: a GfProc over PAR2ProcHIP (parpar_amd/node/gfproc_hip.cc)
driven from JavaScript the way lib/par2.js drives ParPar's GfProc, against the
oracle.  Covers refused adds retried from the progress callback, a slice split
over two GPU backends, chunked passes (setCurrentSliceSize) and partial
inputs."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from parpar_amd import build as BLD

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "js", "gfproc_hip_driver.js")


def say(*a):
    print(*a, flush=True)


def run_job(tmp_path, S, rows, passes, hip, staging, plan=None):
    if shutil.which("node") is None or not os.path.exists(BLD.NODE_ADDON):
        pytest.skip("node or the binding is not available")
    inp_file = tmp_path / "inputs.bin"
    out_file = tmp_path / "outputs.bin"
    inp_file.write_bytes(np.stack(rows).tobytes())
    job = {"slice_size": S, "staging": staging, "hip": hip, "input_file": str(inp_file),
           "output_file": str(out_file), "passes": passes, "plan": plan}
    jf = tmp_path / "job.json"
    jf.write_text(json.dumps(job))
    r = subprocess.run(["node", DRIVER, str(jf)], capture_output=True, text=True, timeout=120)
    say(r.stdout.strip(), r.stderr.strip()[-2000:])
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    data = out_file.read_bytes()
    outs, pos = [], 0
    for p in passes:
        R, size = len(p["exps"]), p["cur_size"]
        rec = np.frombuffer(data, np.uint8, R * size, pos).reshape(R, size)
        pos += R * size
        valid = np.frombuffer(data, np.uint8, R, pos)
        pos += R
        outs.append((rec, valid))
    assert pos == len(data)
    return res, outs


def expected(rows, S, p):
    n = len(p["inputs"])
    inp = np.zeros((n, p["cur_size"]), np.uint8)
    lens = []
    for k, i in enumerate(p["inputs"]):
        inp[k, :i["len"]] = rows[i["row"]][p["offset"]:p["offset"] + i["len"]]
        lens.append(i["len"])
    nums = np.array([i["num"] for i in p["inputs"]], np.uint16)
    return O.recover(inp, lens, p["cur_size"], np.array(p["exps"], np.uint16), input_nums=nums)


def test_node_gfproc_one_gpu(tmp_path):
    S, n = 4104, 13
    rows = [O.prng_bytes(50, 0, k, S) for k in range(n)]
    inputs = [{"row": k, "num": 3 * k + 1, "len": S if k < n - 1 else 1000} for k in range(n)]
    passes = [{"cur_size": S, "offset": 0, "exps": [0, 1, 7, 300, 65534], "inputs": inputs}]
    res, outs = run_job(tmp_path, S, rows, passes, [{"device": 0, "input_batchsize": 4}], staging=2)
    say("one gpu", res)
    rec, valid = outs[0]
    assert (rec == expected(rows, S, passes[0])).all() and valid.all()


def test_node_gfproc_large_slices_one_area(tmp_path):
    """4 MiB slices through one staging area of one input: while a batch's
    copy runs, the next add is refused and retried from the progress callback
    (how often depends on timing, so the count is reported, not asserted)."""
    S, n = 4 << 20, 10
    rows = [O.prng_bytes(52, 0, k, S) for k in range(n)]
    inputs = [{"row": k, "num": k, "len": S} for k in range(n)]
    passes = [{"cur_size": S, "offset": 0, "exps": [1, 4], "inputs": inputs}]
    res, outs = run_job(tmp_path, S, rows, passes, [{"device": 0, "input_batchsize": 1}], staging=1)
    say("large slices: adds refused and retried:", res["refused"])
    rec, valid = outs[0]
    assert (rec == expected(rows, S, passes[0])).all() and valid.all()


def test_node_gfproc_split_and_chunked(tmp_path):
    """Two GPU backends own [0, 2050) and [2050, S) of each slice; two chunked
    passes re-split the current size over them (PAR2Proc::setCurrentSliceSize)."""
    S, n = 6000, 6
    rows = [O.prng_bytes(51, 0, k, S) for k in range(n)]
    hip = [{"device": 0, "slice_offset": 2050, "slice_size": S - 2050, "input_batchsize": 2},
           {"device": 0, "slice_offset": 0, "slice_size": 2050, "input_batchsize": 3, "method": 3}]
    T = 3000
    passes = [{"cur_size": T, "offset": 0, "exps": [2, 5, 9], "inputs": [{"row": k, "num": k, "len": T}
                                                                          for k in range(n)]},
              {"cur_size": S - T, "offset": T, "exps": [2, 5, 9],
               "inputs": [{"row": k, "num": k, "len": S - T if k else 100} for k in range(n)]}]
    res, outs = run_job(tmp_path, S, rows, passes, hip, staging=2)
    say("split", res)
    assert res["devices"] == 2
    for p, (rec, valid) in zip(passes, outs):
        assert (rec == expected(rows, S, p)).all() and valid.all()


def test_node_planned_split(tmp_path):
    """The planner (parpar_amd/node/plan.js) on the visible device: its
    memory comes from hip_device_info, the whole slice goes to it, and a job
    run through the planned specs matches the oracle.  (The planner refuses a
    device listed twice; its arithmetic over several devices is tested on the
    CPU with fixed device infos.)"""
    S, n = 5000, 5
    rows = [O.prng_bytes(53, 0, k, S) for k in range(n)]
    passes = [{"cur_size": S, "offset": 0, "exps": [0, 3, 11],
               "inputs": [{"row": k, "num": 2 * k, "len": S} for k in range(n)]}]
    res, outs = run_job(tmp_path, S, rows, passes, None, staging=2, plan=[{"device": 0, "input_batchsize": 2}])
    say("planned", res["hip"], res["device_info"])
    assert res["hip"] == [{"device": 0, "slice_offset": 0, "slice_size": S, "input_batchsize": 2}]
    assert res["device_info"] and res["device_info"][0]["memory_global"] > (1 << 30)
    assert res["device_info"][0]["compute_units"] > 0
    rec, valid = outs[0]
    assert (rec == expected(rows, S, passes[0])).all() and valid.all()
