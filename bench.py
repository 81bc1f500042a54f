"""Benchmark: GF(2^16) recovery (out_r ^= sum_i c(r,i) * in_i, PAR2 coefficients)
on one MI355X.

Default workload: 2000 input slices x 1 MiB (uniform random, seed 1), 200
recovery slices with exponents 0..199, everything resident in HBM.  One step is
the whole device path over that batch: input preparation (zero padding,
gf16_cksum, log-domain conversion), the plain mul-add, and the gf16_cksum check
of every recovery slice.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  value = GF16 mul-add GB/s of the whole step
(2 bytes per mul-add, MA = inputs x recovery x words); roofline is the plain
mul-add kernel's mul-add-equivalent bandwidth against HBM peak, timed with HIP
events on its own stream; cpu_baseline times the oracle (the C restatement of
PAR2ProcCPU) on a bounded column window of the same workload with every host
thread.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# DRAM bytes per k_muladd_plain launch measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE runs, scaled by a 1 GiB plain copy in the same runs
# (profiles/k_muladd_plain_2000x1MiB_R200.md), keyed by (inputs, slice, recovery).
MEASURED_TRAFFIC = {(2000, 1 << 20, 200): 19.21e9 + 0.210e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--inputs", type=int, default=2000)
    ap.add_argument("--slice", type=int, default=1 << 20)
    ap.add_argument("--recovery", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--cksum", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run one plain device copy of this many bytes (HBM counter reference)")
    return ap.parse_args()


def cpu_baseline(args, gpu_out_window, window_bytes_used):
    """Oracle on the first `window` bytes of every slice (same inputs, same outputs)."""
    import numpy as np
    import oracle as O

    n, R, S = args.inputs, args.recovery, args.slice
    exps = np.arange(R, dtype=np.uint16)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))

    def run(win):
        inp = np.stack([O.prng_bytes(args.seed, 0, i, win) for i in range(n)])
        t0 = time.perf_counter()
        out = O.recover(inp, [win] * n, win, exps, threads=threads)
        return time.perf_counter() - t0, out

    # calibrate on one 32 KiB column tile (one OpenMP work item), then size the
    # window so that `threads` tiles in parallel take about cpu_seconds
    tile = 32768
    t, _ = run(tile)
    tiles = int(args.cpu_seconds * threads / max(t, 1e-9))
    win = max(tile, min(tiles * tile, window_bytes_used, S))
    win -= win % 1024
    log("cpu baseline: %d threads, window %d bytes of each slice" % (threads, win))
    t, out = run(win)
    ma = n * R * (win // 2)
    check = None
    if gpu_out_window is not None:
        w = min(win, gpu_out_window.shape[1])
        check = bool((out[:, :w] == gpu_out_window[:, :w]).all())
    return {
        "value": round(2.0 * ma / t / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": "first %d bytes of all %d input slices x %d recovery slices (%.3g mul-adds, %.1f s)" % (
            win, n, R, ma, t),
        "matches_gpu": check,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs on a single device (byte-range sharding across "
              "GPUs is a design note only); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import parpar_amd
    from parpar_amd import proc as P

    parpar_amd.require_device()
    L = parpar_amd.lib()
    chk = parpar_amd.check
    n, R, S, ck = args.inputs, args.recovery, args.slice, args.cksum
    ck_off, slot = P.slot_layout(S, ck)
    words = S // 2
    ma = n * R * words
    log("workload: %d inputs x %d bytes, %d recovery slices, MA=%.4g" % (n, S, R, ma))

    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    L.pph_device_info(0, name, 256, ctypes.byref(cus))

    stream = ctypes.c_void_p()
    chk(L.pph_stream_create(ctypes.byref(stream)))
    src = P.DeviceBuffer(slot * n)      # data-form inputs (arrive in HBM)
    prepd = P.DeviceBuffer(slot * n)    # log-domain inputs with checksums
    out = P.DeviceBuffer(slot * R)
    valid = P.DeviceBuffer(4 * R)
    d_logs = P.DeviceBuffer(2 * n)
    d_logs.upload(np.array([L.pph_input_log(i) for i in range(n)], np.uint16))
    d_exps = P.DeviceBuffer(2 * R)
    d_exps.upload(np.arange(R, dtype=np.uint16))
    chk(L.pph_fill_random(src.ptr, slot, n, S, args.seed, 0, 0, stream), "fill")
    chk(L.pph_stream_synchronize(stream))
    log("inputs generated in HBM")

    t_step = ctypes.c_void_p()
    t_ker = ctypes.c_void_p()
    chk(L.pph_timer_create(ctypes.byref(t_step)))
    chk(L.pph_timer_create(ctypes.byref(t_ker)))

    def step(timed):
        chk(L.pph_prep_slots_from(prepd.ptr, src.ptr, slot, n, None, S, ck, 1, stream), "prep")
        if timed:
            chk(L.pph_timer_start(t_ker, stream))
        chk(L.pph_muladd_plain(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, None, 0, out.ptr, R, 0, args.group,
                               stream), "muladd")
        if timed:
            chk(L.pph_timer_stop(t_ker, stream))
        chk(L.pph_verify_slots(out.ptr, slot, R, S, ck, valid.ptr, stream), "verify")

    for w in range(args.warmup):
        step(False)
        chk(L.pph_stream_synchronize(stream))
        log("warmup %d done" % (w + 1))

    if args.calib_copy:
        cbytes = args.calib_copy - args.calib_copy % 16
        ca = P.DeviceBuffer(cbytes)
        cb = P.DeviceBuffer(cbytes)
        chk(L.pph_copy(cb.ptr, ca.ptr, cbytes, stream), "copy")
        chk(L.pph_stream_synchronize(stream))
        log("calibration copy of %d bytes done" % cbytes)
        del ca, cb

    ker_ms = []
    chk(L.pph_synchronize())
    t0 = time.perf_counter()
    chk(L.pph_timer_start(t_step, stream))
    for s in range(args.steps):
        step(True)
        ms = ctypes.c_float()
        chk(L.pph_timer_elapsed_ms(t_ker, ctypes.byref(ms)))
        ker_ms.append(ms.value)
    chk(L.pph_timer_stop(t_step, stream))
    chk(L.pph_stream_synchronize(stream))
    chk(L.pph_synchronize())
    wall = time.perf_counter() - t0
    total_ms = ctypes.c_float()
    chk(L.pph_timer_elapsed_ms(t_step, ctypes.byref(total_ms)))
    log("timed %d steps: %.2f ms wall, %.2f ms events" % (args.steps, wall * 1e3, total_ms.value))

    ok = valid.download(4 * R).view(np.int32)
    if not (ok == 1).all():
        raise SystemExit("recovery checksum mismatch on the device: %d of %d slices invalid" % (int((ok != 1).sum()), R))

    ms_per_step = wall * 1e3 / args.steps
    value = 2.0 * ma / (ms_per_step / 1e3) / 1e9
    ker_avg = sum(ker_ms) / len(ker_ms)
    achieved = 2.0 * ma / (ker_avg / 1e3) / 1e9

    # window of the GPU outputs for the CPU cross-check
    win = min(S, 1 << 20)
    gpu_win = None
    cpu = None
    if not args.no_cpu:
        host = out.download(R * slot).reshape(R, slot)
        gpu_win = host[:, :win]
        cpu = cpu_baseline(args, gpu_win, win)
        log("cpu baseline: %s" % cpu)

    res = {
        "metric": "GF16 mul-add GB/s (2 B per GF(2^16) mul-add), plain form",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16 (GF(2^16) words)",
        "data": "synthetic (counter-based random bytes, seed %d)" % args.seed,
        "config": {"workload": "plain recovery: %d inputs x %d B slices, %d recovery slices (exp 0..%d)" % (
            n, S, R, R - 1), "inputs": n, "slice_bytes": S, "recovery": R, "cksum_bytes": ck,
            "mul_adds": ma, "device": name.value.decode(), "cus": cus.value},
        "ma_per_s": round(ma / (ms_per_step / 1e3), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((n, S, R)),
                     "kernel": "k_muladd_plain", "kernel_ms": round(ker_avg, 3),
                     "algorithmic_bytes": 2 * ma},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    L.pph_timer_destroy(t_step)
    L.pph_timer_destroy(t_ker)
    L.pph_stream_destroy(stream)


if __name__ == "__main__":
    main()
