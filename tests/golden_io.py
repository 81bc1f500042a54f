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
