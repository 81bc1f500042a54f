"""Time the reference's own CPU path on the plain benchmark workload and write
tests/golden/reference_timing.json.

Runs only in a container that holds the reference sources: it loads
oracle/_ref/libparpar_ref.so (PAR2Proc + PAR2ProcCPU compiled by
`make -C oracle ref` from the reference's gf16/ sources, GF16_AUTO, every host
thread).  The workload: 2000 input slices of 1 MiB (the counter-based stream,
seed 1), 200 recovery slices with exponents 0..199, timed from the first
addInput to the last getOutput with the inputs already in memory, median of
three runs.  The outputs are compared with the oracle on a column window and
their SHA-256 over all 200 slices is recorded.

    python tests/golden/time_reference.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def main():
    n, S, R, seed = 2000, 1 << 20, 200, 1
    info = O.ref_info()
    inp = np.empty((n, S), np.uint8)
    for i in range(n):
        inp[i] = O.prng_bytes(seed, 0, i, S)
    exps = np.arange(R, dtype=np.uint16)
    times, digest = [], None
    for run in range(3):
        out, valid, secs = O.ref_recover(inp, [S] * n, S, exps)
        assert valid.all()
        times.append(secs)
        d = hashlib.sha256(out.tobytes()).hexdigest()
        assert digest in (None, d)
        digest = d
        print("run %d: %.2f s" % (run + 1, secs), flush=True)
    win = 4096
    ref = O.recover(np.ascontiguousarray(inp[:, :win]), [win] * n, win, exps)
    assert (ref == out[:, :win]).all()
    ma = n * R * (S // 2)
    t = float(np.median(times))
    rec = {"workload": "2000 input slices x 1 MiB (seed 1), 200 recovery slices (exponents 0..199)",
           "method": info["method"], "cksum_size": info["cksum_size"], "threads": info["threads"],
           "host": "the build container (%d vCPU)" % os.cpu_count(),
           "seconds": [round(x, 3) for x in times], "median_seconds": round(t, 3), "mul_adds": ma,
           "ma_per_s": round(ma / t, 1), "gbps_mul_add_equivalent": round(2 * ma / t / 1e9, 3),
           "outputs_sha256": digest}
    with open(os.path.join(HERE, "reference_timing.json"), "w") as f:
        json.dump(rec, f, indent=1)
        f.write("\n")
    print(rec)


if __name__ == "__main__":
    main()
