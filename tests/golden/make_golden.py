"""Write the golden fixtures in tests/golden/ from the reference's own CPU path.

Runs only in a container that holds the reference sources: it needs
oracle/_ref/libparpar_ref.so, built by `make -C oracle ref` from the reference's
gf16/ sources (PAR2Proc + PAR2ProcCPU, GF16_AUTO).  The fixtures are data only:
inputs are named by their synthetic-stream key (seed, job, slice; see
oracle/po_prng.h) and outputs are stored as bytes or digests.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def make_inputs(seed, n, slice_size, lens, job=0, nums=None):
    nums = list(range(n)) if nums is None else nums
    inp = np.zeros((n, max(slice_size, 1)), dtype=np.uint8)
    for k in range(n):
        if lens[k]:
            inp[k, :lens[k]] = O.prng_bytes(seed, job, nums[k], lens[k])
    return inp


def small_cases():
    """Tiny full vectors: slice sizes, input counts, truncation, exponents, paths."""
    cases = []
    cid = 0
    for S in (2, 4, 6, 62, 64, 66, 130, 4096):
        for n in (1, 2, 7, 13):
            for exps in ([1], [0, 1, 65534]):
                lens = [S] * n
                if n > 1 and S > 2:
                    lens[-1] = 1 + (S * 7 + n) % (S - 1)   # truncated last slice, odd allowed
                cases.append(dict(id=cid, seed=100 + cid, n=n, S=S, lens=lens, exps=exps,
                                  nums=list(range(n)), coeffs=None, split=0))
                cid += 1
    # large input numbers and the maximum exponent
    for S, n in ((64, 3), (1030, 5)):
        nums = [32767, 32766, 16384, 1, 0][:n]
        cases.append(dict(id=cid, seed=100 + cid, n=n, S=S, lens=[S] * n, exps=[65534, 32767, 2, 0],
                          nums=nums, coeffs=None, split=0))
        cid += 1
    # explicit coefficients (controller.h:170), including zero and one
    rng = np.random.default_rng(7)
    for S, n, r in ((64, 4, 3), (4096, 9, 5)):
        cf = rng.integers(0, 65536, size=(n, r), dtype=np.uint16)
        cf[0, 0] = 0
        cf[1, 0] = 1
        cases.append(dict(id=cid, seed=100 + cid, n=n, S=S, lens=[S] * n, exps=[1] * r,
                          nums=list(range(n)), coeffs=cf.tolist(), split=0))
        cid += 1
    # byte-range split over two backends (controller.cpp:13-84): even split point,
    # second part odd-sized where the slice is odd
    for S, n, split in ((4096, 6, 1024), (4098, 5, 2050), (1001, 3, 500), (200, 4, 2)):
        lens = [S] * n
        lens[-1] = S // 3
        cases.append(dict(id=cid, seed=100 + cid, n=n, S=S, lens=lens, exps=[0, 1, 2, 500],
                          nums=list(range(n)), coeffs=None, split=split))
        cid += 1
    return cases


def main():
    info = O.ref_info()
    print("reference method:", info)
    manifest = {"reference_method": info["method"], "cksum_size": info["cksum_size"],
                "generator": "tests/golden/make_golden.py (PAR2Proc + PAR2ProcCPU, GF16_AUTO)"}

    # --- PRNG first bytes (pins the synthetic-input generator)
    prng = []
    for key in ((0, 0, 0), (0, 0, 99), (1, 0, 1999), (2, 63, 2999), (3, 0, 32767), (4, 5, 7)):
        prng.append({"key": list(key), "first64": bytes(O.prng_bytes(*key, 64)).hex(),
                     "at_1000003_len13": bytes(O.prng_bytes(*key, 13, offset=1000003)).hex()})
    manifest["prng"] = prng

    # --- coefficients (gfmat_coeff.c), straight from the reference
    R = O.ref()
    ins = [0, 1, 2, 3, 99, 1999, 32767]
    ex = [0, 1, 2, 199, 8191, 65534]
    manifest["input_log"] = {str(i): int(R.ref_gfmat_input_log(i)) for i in ins}
    manifest["coeff"] = {"%d,%d" % (i, e): int(R.ref_gfmat_coeff(i, e)) for i in ins for e in ex}
    all_logs = np.array([R.ref_gfmat_input_log(i) for i in range(32768)], dtype=np.uint16)
    manifest["input_log_sha256"] = hashlib.sha256(all_logs.tobytes()).hexdigest()
    grid_e = [1, 2, 3, 255, 256, 4097, 65534]
    grid = np.array([[R.ref_gfmat_coeff(i, e) for e in grid_e] for i in range(0, 32768, 3)], dtype=np.uint16)
    manifest["coeff_grid"] = {"inputs": "range(0, 32768, 3)", "exps": grid_e,
                              "sha256": hashlib.sha256(grid.tobytes()).hexdigest()}
    manifest["gf16_exp"] = {str(v): int(R.ref_gf16_exp(v)) for v in (0, 1, 7, 8, 15, 16, 255, 65534, 65535 % 65536)}

    # --- gf16_cksum trailers at widths 8/16/32/64 (methods Lookup/SSSE3/AVX2/Affine-AVX512)
    ck = []
    for method in (1, 9, 11, 22):
        width = R.ref_method_cksum_size(method)
        for (S, src) in ((1048576, 1048576), (1048576, 777777), (130, 77), (4096, 4000), (62, 0),
                         (6, 3), (64, 64), (66, 65), (2, 2), (1000, 999)):
            key = (9, method, S * 7 + src)
            data = O.prng_bytes(*key, src) if src else np.zeros(0, np.uint8)
            padded, trailer = O.ref_cksum_copy(data, S, method=method)
            chk_dst = np.zeros(S, np.uint8)
            full = np.concatenate([padded, trailer])
            ok = R.ref_cksum_check(method, O._ptr(chk_dst), O._ptr(full), S)
            ck.append({"method": method, "width": width, "key": list(key), "src_len": src, "slice_len": S,
                       "trailer": bytes(trailer).hex(), "check_ok": int(ok)})
    manifest["cksum"] = ck

    # --- tiny full recovery vectors
    cases = small_cases()
    blobs = {}
    for c in cases:
        inp = make_inputs(c["seed"], c["n"], c["S"], c["lens"], nums=None)
        coeffs = np.array(c["coeffs"], dtype=np.uint16) if c["coeffs"] is not None else None
        out, valid, _ = O.ref_recover(inp, c["lens"], c["S"], c["exps"], input_nums=c["nums"], coeffs=coeffs,
                                      split=c["split"])
        blobs["out_%d" % c["id"]] = out
        c["valid"] = [bool(v) for v in valid]
        c["trailers"] = [bytes(O.ref_cksum_copy(out[r], c["S"])[1]).hex() for r in range(out.shape[0])]
    manifest["small_cases"] = cases
    np.savez_compressed(os.path.join(HERE, "small_outputs.npz"), **blobs)

    # --- 100 x 1 MiB, seed 0, 10 recovery slices (exp 0..9)
    S, n = 1048576, 100
    inp = make_inputs(0, n, S, [S] * n)
    exps = list(range(10))
    times = []
    for _ in range(3):
        out, valid, secs = O.ref_recover(inp, [S] * n, S, exps)
        times.append(secs)
    manifest["config0"] = {
        "seed": 0, "n": n, "slice": S, "exps": exps,
        "valid": [bool(v) for v in valid],
        "sha256": [hashlib.sha256(out[r].tobytes()).hexdigest() for r in range(len(exps))],
        "trailer": [bytes(O.ref_cksum_copy(out[r], S)[1]).hex() for r in range(len(exps))],
        "head256": [bytes(out[r, :256]).hex() for r in range(len(exps))],
        "tail256": [bytes(out[r, -256:]).hex() for r in range(len(exps))],
        "ref_seconds_median": float(np.median(times)), "ref_seconds": times,
        "ref_threads": info["threads"], "ma": n * len(exps) * (S // 2),
    }
    print("config0 reference: %.3f s median (%s threads) -> %.3g MA/s" % (
        np.median(times), info["threads"], n * len(exps) * (S // 2) / np.median(times)))

    # --- per-slice MD5 / CRC32 with zero padding (hasher_base.h:154-165), textbook routines
    hv = []
    for (S, ln) in ((1048576, 1048576), (1048576, 5), (262144, 262143), (64, 0), (100, 55)):
        key = (4, 1, S + ln)
        data = O.prng_bytes(*key, ln) if ln else np.zeros(0, np.uint8)
        padded = np.zeros(S, np.uint8)
        padded[:ln] = data
        hv.append({"key": list(key), "len": ln, "slice": S,
                   "md5": hashlib.md5(padded.tobytes()).hexdigest(),
                   "crc32": "%08x" % (zlib.crc32(padded.tobytes()) & 0xffffffff),
                   "md5_unpadded": hashlib.md5(data.tobytes()).hexdigest()})
    manifest["hash"] = hv

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print("wrote", len(cases), "small cases")


if __name__ == "__main__":
    main()
