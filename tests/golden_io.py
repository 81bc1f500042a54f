"""Loading helpers for the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DIR = os.path.join(HERE, "golden")

_manifest = None
_outputs = None


def manifest():
    global _manifest
    if _manifest is None:
        with open(os.path.join(GOLDEN_DIR, "golden.json")) as f:
            _manifest = json.load(f)
    return _manifest


def small_outputs():
    global _outputs
    if _outputs is None:
        _outputs = dict(np.load(os.path.join(GOLDEN_DIR, "small_outputs.npz"), allow_pickle=False))
    return _outputs


def case_inputs(case):
    """Regenerate a small case's inputs exactly as make_golden.py did."""
    import oracle as O
    n, S = case["n"], case["S"]
    inp = np.zeros((n, max(S, 1)), dtype=np.uint8)
    for k in range(n):
        ln = case["lens"][k]
        if ln:
            inp[k, :ln] = O.prng_bytes(case["seed"], 0, k, ln)
    return inp


def expected_valid(inp, lens, S, exps, nums, width=64):
    """Checksum flags PAR2ProcCPU reports: all true for even slices; for an odd
    slice the stored checksum covers the whole last word (its high byte lies
    past the slice end), so the check of the S-byte slice can fail."""
    import oracle as O
    n_out = len(exps)
    if S % 2 == 0:
        return np.ones(n_out, bool)
    full = np.zeros((inp.shape[0], S + 1), np.uint8)
    full[:, :S] = inp[:, :S]
    out_full = O.recover(full, lens, S + 1, exps, input_nums=nums)
    return np.array([O.cksum_verify(out_full[r, :S], O.cksum(out_full[r], S + 1, width), width)
                     for r in range(n_out)])
