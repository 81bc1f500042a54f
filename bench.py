"""Benchmark: GF(2^16) recovery (out_r ^= sum_i c(r,i) * in_i, PAR2 coefficients)
on one MI355X.

Workloads (--config):
  0 (default)  config 1's line (the headline), with configs 2, 3 and 4 run
               after it and reported under "configs" in the same line.
  1            2000 input slices x 1 MiB (uniform random, seed 1), 200 recovery
               slices with exponents 0..199, plain form.
  2            64 independent jobs (seed 2): inputs U[10,3000], slice size
               log-uniform in [64 KiB, 4 MiB] (multiple of 4), last input
               truncated to a random length, recovery count U[1,500], all in one
               persistent fused launch replayed from a hipGraph, forms and
               grouping chosen by the autotuner.
  3            32768 input slices x 64 KiB (seed 3), 8192 recovery slices with
               exponents 0..8191, bilinear dense form (rank-81 Karatsuba formula:
               81 GEMMs over 0/1 operands on fp4 MFMA; --form dense runs the 16x16
               bit-matrix int8 MFMA kernel instead), with the plain form timed on
               the same data for the ratio.
  4            serving: files arrive as a Poisson process (20/s, seed 4) for a
               bounded window; file size log-uniform in [1 MiB, 8 GiB], slice
               size log-uniform in [256 KiB, 16 MiB] (multiple of 4, raised so
               that a file has at most 32768 slices), recovery count
               max(1, ceil(rho * slices)) with rho U[1%, 30%]; each file's
               slices are its PAR2 inputs.  Per job: per-slice MD5 / CRC-32,
               file MD5 (on host threads, or on the device with
               --md5-host-threads 0), recovery slices and checksum flags back
               in host memory.  Reports p50 / p99 latency (arrival to last
               output) and offered vs achieved throughput.

One step is the whole device path over the batch, everything resident in HBM:
input preparation (zero padding, gf16_cksum, log domain for the plain form),
the mul-add, and the gf16_cksum check of every recovery slice.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 0..4] [--no-cpu]

Prints one JSON line.  value = GF16 mul-add GB/s of the whole step (2 bytes
per mul-add, MA = inputs x recovery x words).  roofline: the mul-add kernel,
timed with HIP events on its stream -- the plain form against the random LDS
lookup rate measured in the same run (one lookup per mul-add; its
mul-add-equivalent rate against HBM peak beside it), the dense forms against
the MFMA peak of their operand type.  cpu_baseline: the reference's own
PAR2ProcCPU (a portable build of its sources, where shipped) on a bounded
sample of the same workload, with the oracle (the C restatement) beside it,
both compared byte for byte with the GPU's output.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from parpar_amd.workloads import batched_jobs, serving_arrivals  # noqa: E402

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA
FP4_PEAK_TFLOPS = 10100.0  # MI355X dense fp4 MFMA

# DRAM bytes per mul-add launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs, scaled by a 1 GiB plain copy in the same runs
# (profiles/k_muladd_plain_2000x1MiB_R200.md), keyed by (form, inputs, slice, recovery).
MEASURED_TRAFFIC = {("plain", 2000, 1 << 20, 200): 27.29e9 + 0.655e9,
                    # k_blm_gemm<256>, the dominant kernel of the bilinear call
                    # (profiles/bilinear_32768x64KiB_R8192.md)
                    ("bilinear", 32768, 1 << 16, 8192): 435.6e9 + 2.77e9}

CONFIGS = {
    # 0: the default run -- the plain workload's headline line, with the batched,
    # dense and serving workloads measured after it and reported in the same line
    0: dict(inputs=2000, slice=1 << 20, recovery=200, seed=1, form="plain", steps=10, warmup=2),
    1: dict(inputs=2000, slice=1 << 20, recovery=200, seed=1, form="plain", steps=10, warmup=2),
    2: dict(inputs=None, slice=None, recovery=None, seed=2, form="batched", steps=2, warmup=1),
    3: dict(inputs=32768, slice=1 << 16, recovery=8192, seed=3, form="bilinear", steps=2, warmup=1),
    4: dict(inputs=None, slice=None, recovery=None, seed=4, form="serving", steps=1, warmup=1),
}


def main_serving(args):
    import threading
    import numpy as np
    import parpar_amd
    from parpar_amd import serve as SV

    parpar_amd.require_device()
    L = parpar_amd.lib()
    chk = parpar_amd.check
    arrivals = serving_arrivals(args.seed, args.window)
    total_bytes = sum(a[1] for a in arrivals)
    total_ma = sum(-(-T // S) * R * (S // 2) for _, T, S, R in arrivals)
    log("serving workload: %d files in %.1f s, %.1f GiB, MA=%.4g" % (
        len(arrivals), args.window, total_bytes / 2**30, total_ma))
    free, tot = ctypes.c_size_t(), ctypes.c_size_t()
    chk(L.pph_mem_get_info(ctypes.byref(free), ctypes.byref(tot)))
    bl_ws = args.bilinear_workspace << 30
    budget = min(free.value - (16 << 30) - bl_ws, 240 << 30)
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    L.pph_device_info(0, name, 256, ctypes.byref(cus))
    pool = None
    if args.host_data:
        # host-resident files: every file is a window of one host pool (random bytes from the
        # device generator), read through the pinned staging path and hashed on host threads
        pool_bytes = max(args.pool_gib << 30, max(a[1] for a in arrivals))
        pool = np.empty(pool_bytes, np.uint8)
        chunk = 1 << 30
        buf = parpar_amd.proc.DeviceBuffer(chunk)
        for c in range(0, pool_bytes, chunk):
            n = min(chunk, pool_bytes - c)
            chk(L.pph_fill_random(buf.ptr, n, 1, n, args.seed, 0xD0D0, c // chunk, None))
            chk(L.pph_synchronize())
            chk(L.pph_memcpy_d2h(parpar_amd.proc._ptr(pool[c:c + n]), buf.ptr, n))
        buf.close()
        log("host pool %.1f GiB" % (pool_bytes / 2**30))

    def pool_off(k, T):
        return (k * 0x9E3779B1 * 4096 + 12345) % (len(pool) - T + 1)

    srv = SV.Server(budget, max_compute=args.max_compute, bilinear_workspace=bl_ws,
                    md5_host_threads=args.md5_host_threads,
                    upload_threads=args.upload_threads if args.host_data else 0)
    log("arena %.1f GiB, %d compute streams, bilinear workspace %d GiB, %d file-MD5 host threads%s" % (
        budget / 2**30, args.max_compute, args.bilinear_workspace, args.md5_host_threads,
        ", host data: %d upload threads" % args.upload_threads if args.host_data else ""))
    # warm-up job (tables, first launches), not counted
    srv.run_all([SV.FileJob(8 << 20, 1 << 20, 4, seed=args.seed, job_id=1 << 30)])

    # which jobs the CPU checks: full hashes for small files, recovery windows for a spread
    small = [k for k, a in enumerate(arrivals) if a[1] <= (64 << 20)][:12]
    by_size = sorted(range(len(arrivals)), key=lambda k: arrivals[k][1])
    rec_sample = sorted(set(small[:6] + by_size[len(by_size) // 2:len(by_size) // 2 + 3] + by_size[-2:]))
    keep = set(small) | set(rec_sample)

    jobs = [SV.FileJob(T, S, R, seed=args.seed, job_id=k,
                       data=None if pool is None else pool[pool_off(k, T):pool_off(k, T) + T])
            for k, (_, T, S, R) in enumerate(arrivals)]
    t0 = time.perf_counter()

    def submitter():
        for k, a in enumerate(arrivals):
            d = t0 + a[0] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(jobs[k])
            jobs[k] = None if k not in keep else jobs[k]

    th = threading.Thread(target=submitter)
    th.start()
    results = []
    records = []
    kept = {}
    last_print = t0
    while len(results) < len(arrivals):
        for job in srv.poll(wait_ms=500):
            results.append(job.result)
            records.append(dict(job_id=job.job_id, file_size=job.file_size, slice_size=job.slice_size,
                                n_in=job.n_in, n_out=job.n_out, **{k: round(job.result[k], 4) for k in (
                                    "latency_s", "queue_s", "compute_s", "run_s")}))
            if job.job_id in keep:
                kept[job.job_id] = job
            if job.result["status"] != 0:
                raise SystemExit("serving job %d failed: %s" % (job.job_id, job.result))
        now = time.perf_counter()
        if now - last_print > 10:
            st = srv.stats()
            log("t=%.0f s: %d/%d done, %d in flight, arena peak %.1f GiB, %d hash rounds" % (
                now - t0, len(results), len(arrivals), st["submitted"] - st["completed"],
                st["peak_bytes"] / 2**30, st["md5_rounds"]))
            last_print = now
    makespan = time.perf_counter() - t0
    th.join()
    st = srv.stats()
    srv.close()
    lat = np.array([r["latency_s"] for r in results])
    comp = np.array([r["compute_s"] + r["queue_s"] for r in results])
    bad_flags = sum(int((kept[k].valid[:kept[k].n_out] != 1).sum()) for k in kept)
    log("done in %.1f s; latency p50 %.2f s p99 %.2f s" % (makespan, np.percentile(lat, 50), np.percentile(lat, 99)))
    if args.dump_jobs:
        with open(args.dump_jobs, "w") as f:
            for r in records:
                f.write(json.dumps(r) + "\n")

    cpu = None
    if not args.no_cpu:
        import hashlib
        import zlib
        import oracle as O
        threads = cpu_threads()
        ok = bad_flags == 0
        t_cpu, cma = 0.0, 0
        for k in small:
            job = kept[k]
            if pool is not None:
                data = job.data
            else:
                data = np.concatenate([O.prng_bytes(args.seed, k, i,
                                                    min(job.slice_size, job.file_size - i * job.slice_size))
                                       for i in range(job.n_in)])
            ok = ok and bytes(job.file_md5).hex() == hashlib.md5(data.tobytes()).hexdigest()
            pad = np.zeros(job.n_in * job.slice_size, np.uint8)
            pad[:job.file_size] = data
            pad = pad.reshape(job.n_in, job.slice_size)
            for i in range(job.n_in):
                row = pad[i].tobytes()
                ok = ok and bytes(job.slice_md5[i]).hex() == hashlib.md5(row).hexdigest()
                ok = ok and int(job.slice_crc[i]) == zlib.crc32(row) & 0xffffffff
        rng = np.random.default_rng(1000 + args.seed)
        for k in rec_sample:
            job = kept[k]
            S, n = job.slice_size, job.n_in
            win = min(4096, S)
            off = int(rng.integers(0, S - win + 1)) & ~1
            last = job.file_size - (n - 1) * S
            lens = [win] * n
            lens[-1] = max(0, min(win, last - off))
            inp = np.zeros((n, win), np.uint8)
            for i in range(n):
                if lens[i] and pool is not None:
                    inp[i, :lens[i]] = job.data[i * S + off:i * S + off + lens[i]]
                elif lens[i]:
                    inp[i, :lens[i]] = O.prng_bytes(args.seed, k, i, lens[i], offset=off)
            tc = time.perf_counter()
            ref = O.recover(inp, lens, win, np.arange(job.n_out, dtype=np.uint16), threads=threads)
            t_cpu += time.perf_counter() - tc
            cma += n * job.n_out * (win // 2)
            ok = ok and bool((job.recovery[:, off:off + win] == ref).all())
        cpu = {"value": round(2.0 * cma / t_cpu / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": "4 KiB column windows of %d jobs' recovery (%.3g mul-adds, %.1f s); full per-slice and "
                         "file hashes of %d small files checked with hashlib / zlib" % (
                             len(rec_sample), cma, t_cpu, len(small)),
               "matches_gpu": ok}
        log("cpu check: %s" % cpu)

    achieved = 2.0 * total_ma / makespan / 1e9
    res = {
        "metric": "GF16 mul-add GB/s (2 B per GF(2^16) mul-add), serving under Poisson arrivals",
        "value": round(achieved, 2), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(makespan * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u16 (GF(2^16) words); MD5 / CRC-32 bytes",
        "data": ("synthetic (counter-based random bytes generated on the device, seed %d); %s" % (
                     args.seed, "file MD5 on %d host threads, read back from HBM" % args.md5_host_threads
                     if args.md5_host_threads else "file MD5 in shared device rounds")
                 if pool is None else
                 "synthetic, host-resident: each file a window of a %.0f GiB host pool of random bytes, "
                 "copied through pinned staging by %d threads; file MD5 on %d host threads" % (
                     len(pool) / 2**30, args.upload_threads, args.md5_host_threads)),
        "config": {"workload": "serving: Poisson 20 files/s for %.0f s, file 1 MiB..8 GiB, slice 256 KiB..16 MiB, "
                               "recovery 1..30%% of slices" % args.window, "config_index": 4,
                   "files": len(arrivals), "file_bytes": total_bytes, "mul_adds": total_ma,
                   "arena_bytes": budget, "compute_streams": args.max_compute,
                   "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default"), "bilinear_workspace": bl_ws,
                   "device": name.value.decode(), "cus": cus.value},
        "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
        "latency_p99_s": round(float(np.percentile(lat, 99)), 3),
        "latency_max_s": round(float(lat.max()), 3),
        "recovery_latency_p50_s": round(float(np.percentile(comp, 50)), 3),
        "recovery_latency_p99_s": round(float(np.percentile(comp, 99)), 3),
        "offered": {"bytes_per_s": round(total_bytes / args.window / 1e9, 3), "ma_per_s": total_ma / args.window,
                    "unit": "GB/s of file bytes"},
        "achieved": {"bytes_per_s": round(total_bytes / makespan / 1e9, 3), "ma_per_s": total_ma / makespan,
                     "makespan_s": round(makespan, 2)},
        "engine": {"peak_jobs": st["peak_jobs"], "peak_arena_bytes": st["peak_bytes"], "md5_rounds": st["md5_rounds"],
                   "bilinear_jobs": st["bilinear_jobs"], "host_md5_jobs": st["host_md5_jobs"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole serving pipeline (mul-add-equivalent bytes over the makespan)"},
        "cpu_baseline": cpu,
    }
    return res


def main_backend(args):
    """The plain workload through the drop-in boundary: PAR2ProcHIP (pph_proc_*)
    fed host buffers the way ParPar's controller does (addInput per slice,
    endInput, getOutput per recovery slice), so PCIe both ways is inside the
    timed region.  Not the headline line: the device-resident rate is."""
    import numpy as np
    import parpar_amd
    from parpar_amd import proc as P
    import oracle as O

    parpar_amd.require_device()
    L = parpar_amd.lib()
    chk = parpar_amd.check
    n, R, S = args.inputs, args.recovery, args.slice
    ma = n * R * (S // 2)
    log("backend workload: %d inputs x %d B, %d recovery slices, MA=%.4g (host buffers)" % (n, S, R, ma))
    # host inputs: the device generator's bytes, copied down once
    inputs = np.empty((n, S), np.uint8)
    chunk = max(1, (256 << 20) // S)
    buf = P.DeviceBuffer(chunk * S)
    for k0 in range(0, n, chunk):
        c = min(chunk, n - k0)
        chk(L.pph_fill_random(buf.ptr, S, c, S, args.seed, 0, k0, None))
        chk(L.pph_synchronize())
        chk(L.pph_memcpy_d2h(P._ptr(inputs[k0:k0 + c]), buf.ptr, c * S))
    buf.close()
    out = np.zeros((R, S), np.uint8)
    be = P.PAR2ProcHIP(0, 2)
    be.setSliceSize(S)
    forms = {"auto": parpar_amd.FORM_AUTO, "plain": parpar_amd.FORM_PLAIN, "bilinear": parpar_amd.FORM_BILINEAR}
    be.init(args.cksum, args.batch, forms[args.backend_form])
    in_batch = be.getInputBatchSize()  # the backend's default when --batch is 0
    be.setRecoverySlices(np.arange(R, dtype=np.uint16))

    phases = np.zeros(3)  # seconds in the adds, endInput, the getOutputs

    def step():
        t = [time.perf_counter()]
        be.discardOutput()
        for i in range(n):
            if be.canAdd() == parpar_amd.PPH_ADD_FULL:
                be.waitForAdd()
            be.addInput(inputs[i], S, inputNum=i)
        t.append(time.perf_counter())
        be.endInput()
        t.append(time.perf_counter())
        ok = True
        for r in range(R):
            ok = be.getOutput(r, out[r]) and ok
        t.append(time.perf_counter())
        phases[:] += np.diff(t)
        return ok

    for _ in range(args.warmup):
        step()
    phases[:] = 0
    times, valid = [], True
    for _ in range(args.steps):
        t0 = time.perf_counter()
        valid = step() and valid
        times.append(time.perf_counter() - t0)
    be.close()
    ms = 1e3 * float(np.mean(times))
    rate = 2.0 * ma / (ms / 1e3) / 1e9
    phase_ms = [round(1e3 * x / args.steps, 2) for x in phases]
    log("backend: %.1f ms per pass (adds %.1f, endInput %.1f, getOutputs %.1f), %.1f GB/s mul-add-equivalent, "
        "flags %s" % (ms, *phase_ms, rate, valid))
    # CPU check on a column window of every input and recovery slice
    win = min(4096, S)
    tc = time.perf_counter()
    ref = O.recover(np.ascontiguousarray(inputs[:, :win]), [win] * n, win, np.arange(R, dtype=np.uint16),
                    threads=cpu_threads())
    t_cpu = time.perf_counter() - tc
    ok = bool((out[:, :win] == ref).all()) and valid
    res = {
        "metric": "GF16 mul-add GB/s (2 B per GF(2^16) mul-add) through the PAR2ProcHIP boundary, host buffers",
        "value": round(rate, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u16 (GF(2^16) words)",
        "data": "synthetic (counter-based random bytes, seed %d), in host memory" % args.seed,
        "config": {"workload": "backend pass: %d host inputs x %d B -> %d recovery slices (exp 0..%d), input batch "
                               "%d, addInput/endInput/getOutput with PCIe both ways" % (
                                   n, S, R, R - 1, in_batch), "inputs": n, "slice_bytes": S, "recovery": R,
                   "backend_form": args.backend_form},
        "phases_ms": {"add_inputs": phase_ms[0], "end_input": phase_ms[1], "get_outputs": phase_ms[2]},
        "cpu_check": {"window_bytes": win, "matches_gpu": ok, "mul_adds": n * R * (win // 2),
                      "seconds": round(t_cpu, 2)},
    }
    print(json.dumps(res), flush=True)
    if not ok:
        raise SystemExit("backend output differs from the oracle")


def main_hashes(args):
    """Per-slice MD5 + CRC-32 (pph_slice_hash) and recovery-packet MD5
    (pph_packet_md5_*) on slices generated in HBM, timed with HIP events:
    few large slices (bound by one MD5 chain per slice) to many small ones."""
    import numpy as np
    import parpar_amd
    from parpar_amd import proc as P

    parpar_amd.require_device()
    L = parpar_amd.lib()
    chk = parpar_amd.check
    stream = ctypes.c_void_p()
    chk(L.pph_stream_create(ctypes.byref(stream)))
    t = ctypes.c_void_p()
    chk(L.pph_timer_create(ctypes.byref(t)))
    reps = max(1, args.steps or 3)

    def timed(fn):
        fn()
        chk(L.pph_stream_synchronize(stream))
        best = None
        for _ in range(reps):
            chk(L.pph_timer_start(t, stream))
            fn()
            chk(L.pph_timer_stop(t, stream))
            chk(L.pph_stream_synchronize(stream))
            ms = ctypes.c_float()
            chk(L.pph_timer_elapsed_ms(t, ctypes.byref(ms)))
            best = ms.value if best is None else min(best, ms.value)
        return best

    rows = []
    for count, S in [(64, 16 << 20), (100, 16 << 20), (1024, 1 << 20), (32768, 65536)]:
        _, slot = P.slot_layout(S, 64)
        buf = P.DeviceBuffer(slot * count)
        md5, crc = P.DeviceBuffer(16 * count), P.DeviceBuffer(4 * count)
        chk(L.pph_fill_random(buf.ptr, slot, count, S, 7, 0, 0, stream))
        ms = timed(lambda: chk(L.pph_slice_hash(buf.ptr, slot, count, None, S, md5.ptr, crc.ptr, stream)))
        rows.append({"kernel": "k_slice_hash", "slices": count, "slice_bytes": S, "ms": round(ms, 3),
                     "per_chain_MBps": round(S / ms / 1e3, 1), "GBps": round(count * S / ms / 1e6, 1)})
        log("slice hash %5d x %8d: %.2f ms" % (count, S, ms))
        del buf, md5, crc
    # recovery packets: 200 slices of 1 MiB (the default workload's recovery count and slice size)
    R, S = 200, 1 << 20
    _, slot = P.slot_layout(S, 64)
    buf = P.DeviceBuffer(slot * R)
    chk(L.pph_fill_random(buf.ptr, slot, R, S, 8, 0, 0, stream))
    h = ctypes.c_void_p()
    chk(L.pph_packet_md5_create(R, ctypes.byref(h)))
    dig = P.DeviceBuffer(16 * R)
    sid = np.arange(16, dtype=np.uint8)
    exps = np.arange(R, dtype=np.uint16)

    def packets():
        chk(L.pph_packet_md5_begin(h, sid.ctypes.data, exps.ctypes.data, stream))
        chk(L.pph_packet_md5_update(h, buf.ptr, slot, S, stream))
        chk(L.pph_packet_md5_final(h, dig.ptr, stream))

    ms = timed(packets)
    rows.append({"kernel": "k_packet_md5", "slices": R, "slice_bytes": S, "ms": round(ms, 3),
                 "per_chain_MBps": round(S / ms / 1e3, 1), "GBps": round(R * S / ms / 1e6, 1)})
    log("packet md5 %d x %d: %.2f ms" % (R, S, ms))
    L.pph_packet_md5_destroy(h)
    print(json.dumps({"metric": "hash kernels: best of %d launches" % reps, "unit": "ms", "n_gpus": 1,
                      "data": "synthetic (counter-based random bytes generated on the device)",
                      "rows": rows}))



def batch_row_split(n_out):
    """Padded rows of a bilinear job in a batch: tiles of 256, 128 and 64 rows,
    the cheapest split with 128-row tiles weighed 1.05 and 64-row tiles 1.35
    (the library's blm_row_split)."""
    best = None
    for a in range(-(-n_out // 256) + 1):
        for b in range(4):
            for c in range(4):
                rows = 256 * a + 128 * b + 64 * c
                if rows < n_out:
                    continue
                cost = 256.0 * a + 1.05 * 128 * b + 1.35 * 64 * c + 1e-6 * (a + b + c)
                if best is None or cost < best[0]:
                    best = (cost, rows)
    return best[1]

def main_batched(args):
    import numpy as np
    import parpar_amd
    from parpar_amd import batch as B

    parpar_amd.require_device()
    L = parpar_amd.lib()
    chk = parpar_amd.check
    shapes = batched_jobs(args.seed)
    jobs = []
    ma = 0
    for j, (n, S, last, R) in enumerate(shapes):
        lens = [S] * n
        lens[-1] = last
        jobs.append(B.Job(n, S, np.arange(R), lens=lens, cksum=args.cksum))
        ma += n * R * (S // 2)
    in_bytes = sum(j.slot * j.n_in for j in jobs)
    out_bytes = sum(j.slot * j.n_out for j in jobs)
    log("batched workload: %d jobs, %.1f GiB inputs, %.1f GiB outputs, MA=%.4g" % (
        len(jobs), in_bytes / 2**30, out_bytes / 2**30, ma))
    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    L.pph_device_info(0, name, 256, ctypes.byref(cus))
    stream = ctypes.c_void_p()
    chk(L.pph_stream_create(ctypes.byref(stream)))

    t0 = time.perf_counter()
    tuning = B.autotune()
    log("autotune: %.2f s %s" % (time.perf_counter() - t0, tuning.as_dict()))
    batch = B.Batch(jobs, cksum=args.cksum, tuning=tuning)
    forms = [batch.plan(j)["form"] for j in range(len(jobs))]
    group = batch.plan(0)["group"]
    log("plan: group %d, forms %s" % (group, {f: forms.count(f) for f in set(forms)}))

    def refill():
        for j, job in enumerate(jobs):
            job.fill(args.seed, j, stream)
        chk(L.pph_stream_synchronize(stream))

    # eager run: fused kernel time from HIP events
    refill()
    batch.run(use_graph=False, stream=stream)
    fused_ms = batch.fused_ms()
    chk(L.pph_stream_synchronize(stream))
    log("eager run: fused mul-add %.2f ms" % fused_ms)
    for w in range(args.warmup):
        refill()
        batch.run(use_graph=True, stream=stream)
        chk(L.pph_stream_synchronize(stream))
        log("graph warmup %d done" % (w + 1))
    times = []
    for s_ in range(args.steps):
        refill()
        chk(L.pph_synchronize())
        t0 = time.perf_counter()
        batch.run(use_graph=True, stream=stream)
        chk(L.pph_stream_synchronize(stream))
        times.append(time.perf_counter() - t0)
        log("graph step %d: %.2f ms" % (s_ + 1, times[-1] * 1e3))
    bad = sum(int((job.valid() != 1).sum()) for job in jobs if job.slice_size % 2 == 0)
    if bad:
        raise SystemExit("batched: %d recovery slices failed the device checksum" % bad)
    ms_per_step = 1e3 * sum(times) / len(times)

    cpu = None
    if not args.no_cpu:
        import oracle as O
        threads = cpu_threads()
        rng = np.random.default_rng(1000 + args.seed)
        ok = True
        cma = 0
        t_cpu = 0.0
        for j, job in enumerate(jobs):
            win = min(4096, job.slice_size)
            off = int(rng.integers(0, job.slice_size - win + 1)) & ~1
            lens = [max(0, min(win, int(job.lens[k]) - off)) for k in range(job.n_in)]
            inp = np.zeros((job.n_in, win), np.uint8)
            for k in range(job.n_in):
                if lens[k]:
                    inp[k, :lens[k]] = O.prng_bytes(args.seed, j, k, lens[k], offset=off)
            t0 = time.perf_counter()
            ref = O.recover(inp, lens, win, job.exps, threads=threads)
            t_cpu += time.perf_counter() - t0
            cma += job.n_in * job.n_out * (win // 2)
            got = np.stack([job.d_out.download(win, off=r * job.slot + off) for r in range(job.n_out)])
            ok = ok and bool((got == ref).all())
        cpu = {"value": round(2.0 * cma / t_cpu / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": "a 4 KiB column window of every job (all inputs, all recovery slices; %.3g mul-adds, "
                         "%.1f s)" % (cma, t_cpu), "matches_gpu": ok}
        log("cpu baseline: %s" % cpu)

    # Roofline of the bilinear jobs (all but the smallest; most of the mul-adds):
    # their fp4 MFMA work, 81 components x 2 ops per mul-add, over the fused
    # region's HIP-event time (which also holds the word-plane, recombination,
    # the bilinear jobs' preparation and verification, and the plain and dense
    # launches).  executed_mfma_ops adds the GEMM's padding
    # (inputs to 256, at least 512; outputs to the job's row tiles of 256, 128
    # and 64 rows, split as the library splits them).
    bl_ops = bl_exec = 0
    for k, j in enumerate(jobs):
        if forms[k] != "bilinear":
            continue
        words = j.slot // 2
        kw = max(16, -(-j.n_in // 256) * 8)
        bl_ops += 2 * 81 * j.n_in * j.n_out * words
        bl_exec += 2 * 81 * (32 * kw) * batch_row_split(j.n_out) * words
    achieved_tf = bl_ops / (fused_ms / 1e3) / 1e12
    hbm_eq = 2.0 * ma / (fused_ms / 1e3) / 1e9
    res = {
        "metric": "GF16 mul-add GB/s (2 B per GF(2^16) mul-add), batched uneven jobs",
        "value": round(2.0 * ma / (ms_per_step / 1e3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16 (GF(2^16) words)",
        "data": "synthetic (counter-based random bytes, seed %d)" % args.seed,
        "config": {"workload": "batched: 64 uneven PAR2 jobs in one hipGraph (plain jobs in one persistent "
                               "launch, bilinear jobs in cross-job waves)", "config_index": 2,
                   "jobs": len(jobs), "mul_adds": ma, "input_bytes": in_bytes, "output_bytes": out_bytes,
                   "plain_group": group, "forms": {f: forms.count(f) for f in set(forms)},
                   "device": name.value.decode(), "cus": cus.value},
        "ma_per_s": round(ma / (ms_per_step / 1e3), 1),
        "roofline": {"bound": "mfma", "achieved": round(achieved_tf, 2), "peak": FP4_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tf / FP4_PEAK_TFLOPS, 4), "traffic": None,
                     "kernel": "k_blm_gemm over the bilinear jobs (fused region: + word planes, recombination, "
                               "the bilinear jobs' preparation and verification, k_muladd_batch)",
                     "kernel_ms": round(fused_ms, 3), "algorithmic_mfma_ops": bl_ops,
                     "executed_mfma_ops": bl_exec, "operand_type": "fp4 e2m1",
                     "hbm_equivalent": {"achieved": round(hbm_eq, 2), "unit": "GB/s",
                                        "note": "2 B per mul-add over the fused region, not memory traffic"}},
        "cpu_baseline": cpu,
        "autotune": tuning.as_dict(),
        "jobs": [[j.n_in, j.slice_size, j.n_out, forms[k][0]] for k, j in enumerate(jobs)],
    }
    batch.close()
    del jobs
    L.pph_stream_destroy(stream)
    L.pph_bilinear_release()
    return res


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def config_args(args, config, **over):
    import copy
    b = copy.copy(args)
    b.config = config
    for k, v in CONFIGS[config].items():
        setattr(b, k, v)
    for k, v in over.items():
        setattr(b, k, v)
    return b


def main_all(args):
    """The default run: config 1 (plain form) is the headline; configs 2
    (batched), 3 (bilinear and int8 bit-matrix forms against the plain one) and
    4 (serving, file MD5 on host threads and on the device) follow, each with
    its own CPU check, inside a time budget."""
    t0 = time.perf_counter()

    def left():
        return args.budget_s - (time.perf_counter() - t0)

    log("== config 1 (headline)")
    res = main_kernel(config_args(args, 1, steps=args.steps, warmup=args.warmup))
    res["config"]["config_index"] = 1
    extra = {}

    def run(name, est_s, fn):
        if left() < est_s:
            extra[name] = {"skipped": "time budget: %.0f s left of %.0f, about %.0f s needed" % (
                left(), args.budget_s, est_s)}
            log("== %s skipped (%.0f s left)" % (name, left()))
            return
        log("== %s" % name)
        ts = time.perf_counter()
        extra[name] = fn()
        extra[name]["bench_seconds"] = round(time.perf_counter() - ts, 1)

    run("2_batched", 45, lambda: main_batched(config_args(args, 2, no_cpu=args.no_cpu)))

    def config3():
        r = main_kernel(config_args(args, 3, no_cpu=args.no_cpu))
        d = main_kernel(config_args(args, 3, form="dense", steps=1, warmup=1, no_cpu=True, no_compare=True))
        r["dense_int8"] = {"roofline": d["roofline"], "ma_per_s": d["ma_per_s"], "ms_per_step": d["ms_per_step"],
                           "vs_plain": round(d["ma_per_s"] / r["compare"]["plain_ma_per_s"], 3)
                           if r.get("compare") else None}
        return r

    run("3_dense", 40, config3)
    arrivals = serving_arrivals(4, args.serve_window)
    run("4_serving_file_md5_host", 40,
        lambda: main_serving(config_args(args, 4, window=args.serve_window, md5_host_threads=12, host_data=False,
                                         no_cpu=args.no_cpu)))
    # one wavefront's MD5 chain runs at about 90 MB/s: the largest file bounds this run
    est = max(a[1] for a in arrivals) / 85e6 + 30
    run("4_serving_file_md5_device", est,
        lambda: main_serving(config_args(args, 4, window=args.serve_window, md5_host_threads=0, host_data=False,
                                         no_cpu=args.no_cpu)))
    res["configs"] = extra
    res["bench_seconds"] = round(time.perf_counter() - t0, 1)
    print(json.dumps(res), flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", type=int, default=0, choices=sorted(CONFIGS),
                    help="0 (default): config 1's headline line with configs 2, 3 and 4 measured after it and "
                         "reported in the same line; 1-4: that workload alone")
    ap.add_argument("--budget-s", type=float, default=330.0,
                    help="default run: a later workload is skipped (and says so) when it would end past this many "
                         "seconds from the start")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--inputs", type=int, default=None)
    ap.add_argument("--slice", type=int, default=None)
    ap.add_argument("--recovery", type=int, default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--form", choices=["plain", "dense", "bilinear", "batched", "serving", "backend", "hashes"], default=None)
    ap.add_argument("--batch", type=int, default=0, help="backend: input batch size (0 = the backend's default)")
    ap.add_argument("--backend-form", choices=["auto", "plain", "bilinear"], default="auto",
                    help="backend: kernel form (auto: bilinear for batches of 256 or more inputs with 64 or more "
                         "recovery slices, else plain)")
    ap.add_argument("--window", type=float, default=10.0, help="serving: seconds of arrivals")
    ap.add_argument("--serve-window", type=float, default=3.0,
                    help="default run: seconds of arrivals of its serving workload")
    ap.add_argument("--max-compute", type=int, default=8, help="serving: jobs computing at once")
    ap.add_argument("--bilinear-workspace", type=int, default=40,
                    help="serving: GiB of device workspace for the bilinear form (0: plain only)")
    ap.add_argument("--host-data", action="store_true",
                    help="serving: files in host memory (pinned staging uploads, file MD5 on host threads)")
    ap.add_argument("--pool-gib", type=int, default=16, help="serving --host-data: host pool size")
    ap.add_argument("--md5-host-threads", type=int, default=12,
                    help="serving: host threads for the file MD5 (0: shared hash rounds on the device)")
    ap.add_argument("--upload-threads", type=int, default=4)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="serving: GPU_MAX_HW_QUEUES for this process (>= max-compute + 1 gives every compute "
                         "stream and the hash stream a hardware queue of its own; 0 leaves the environment as is)")
    ap.add_argument("--dump-jobs", default=None, help="serving: write one JSON line per job (sizes, latencies)")
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--cksum", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="dense: skip timing the plain form beside it")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run one plain device copy of this many bytes (HBM counter reference)")
    a = ap.parse_args()
    for k, v in CONFIGS[a.config].items():
        if getattr(a, k) is None:
            setattr(a, k, v)
    return a


def host_cpus():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()


def cpu_threads():
    # the GPU host gives one GPU's job a share of 16 cores, whatever the host has
    return max(1, min(host_cpus(), 16))


def plain_roofline(L, ma, ker_ms, traffic, compulsory):
    """The plain kernel against its bound: one random 16-bit lookup into the
    128 KiB LDS exp table per mul-add, against the lookup rate measured in
    this run by a kernel that does only such lookups (pph_lds_gather_ceiling);
    beside it the mul-add-equivalent rate (2 B per mul-add, the unit of the
    reference's charts) against HBM peak, and the DRAM traffic of a launch."""
    rate = ctypes.c_double()
    cms = ctypes.c_float()
    parpar_amd_check(L.pph_lds_gather_ceiling(8000, ctypes.byref(rate), ctypes.byref(cms)), "lds ceiling")
    achieved = ma / (ker_ms / 1e3) / 1e9
    peak = rate.value / 1e9
    eq = 2.0 * ma / (ker_ms / 1e3) / 1e9
    roof = {"bound": "lds", "achieved": round(achieved, 1), "peak": round(peak, 1), "unit": "Glookups/s",
            "frac": round(achieved / peak, 4), "traffic": traffic,
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of this launch shape, scaled by a plain copy "
                              "(profiles/k_muladd_plain_2000x1MiB_R200.md); not measured in this run",
            "kernel": "k_muladd_plain", "kernel_ms": round(ker_ms, 3), "lookups_per_launch": ma,
            "peak_source": "measured in this run: k_lds_gather_ceiling, random 16-bit lookups into the same "
                           "128 KiB LDS table, one 1024-thread workgroup per CU (%.2f ms launch)" % cms.value,
            "hbm_equivalent": {"achieved": round(eq, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(eq / HBM_PEAK_GBS, 4),
                               "note": "2 B per mul-add (region bytes multiplied), not DRAM traffic"}}
    if traffic:
        roof["dram"] = {"GBps": round(traffic / (ker_ms / 1e3) / 1e9, 1),
                        "frac_of_hbm": round(traffic / (ker_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                        "compulsory_bytes": compulsory,
                        "note": "compulsory: every input slot read once, every recovery slot written once"}
    return roof


def parpar_amd_check(st, what):
    import parpar_amd
    return parpar_amd.check(st, what)


def reference_container_timing():
    """The reference CPU path timed on the full plain workload in the build
    container (tests/golden/reference_timing.json), when present."""
    path = os.path.join(ROOT, "tests", "golden", "reference_timing.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def cpu_baseline_reference(args, gpu_out):
    """The reference's own PAR2Proc + PAR2ProcCPU (portable build of its
    sources, the method its CPUID dispatch picks on this host) on the first
    `win` bytes of every input and recovery slice, sized to about
    --cpu-seconds, against the GPU's bytes; the oracle (the C restatement)
    beside it on a smaller window.  Without the reference build, the oracle
    alone is the baseline."""
    import numpy as np
    import oracle as O

    n, R, S = args.inputs, args.recovery, args.slice
    exps = np.arange(R, dtype=np.uint16)
    threads = cpu_threads()
    inp_full = None

    def inputs(win):
        nonlocal inp_full
        if inp_full is None or inp_full.shape[1] < win:
            inp_full = np.stack([O.prng_bytes(args.seed, 0, i, win) for i in range(n)])
        return np.ascontiguousarray(inp_full[:, :win])

    def sized(run):
        # calibrate on a 32 KiB column window, then size the window to about cpu_seconds
        t, _ = run(32768)
        win = int(args.cpu_seconds / max(t, 1e-9) * 32768)
        win = max(32768, min(win - win % 4096, S))
        return win, run(win)

    port = None
    if O.ref_available(portable=True):
        info = O.ref_info(portable=True)

        def run_ref(win):
            out, valid, secs = O.ref_recover(inputs(win), [win] * n, win, exps, threads=threads, portable=True)
            return secs, (out, valid)

        win, (t, (out, valid)) = sized(run_ref)
        ma = n * R * (win // 2)
        cpu = {"value": round(2.0 * ma / t / 1e9, 3), "unit": "GB/s", "cores": threads, "host_cpus": host_cpus(),
               "kind": "reference", "method": info["method"],
               "sample": "the reference PAR2Proc + PAR2ProcCPU (its own sources, portable build, GF16_AUTO) on the "
                         "first %d bytes of all %d input slices x %d recovery slices (%.3g mul-adds, %.1f s, "
                         "first addInput to last getOutput)" % (win, n, R, ma, t),
               "matches_gpu": bool((out == gpu_out[:, :win]).all()) and bool(valid.all())}
        pwin = max(4096, min(S, win // 4) // 4096 * 4096)
    else:
        cpu = None
        pwin = None

    def run_port(win):
        x = inputs(win)
        t0 = time.perf_counter()
        out = O.recover(x, [win] * n, win, exps, threads=threads)
        return time.perf_counter() - t0, out

    if pwin is None:
        pwin, (t, out) = sized(run_port)
    else:
        t, out = run_port(pwin)
    ma = n * R * (pwin // 2)
    port = {"value": round(2.0 * ma / t / 1e9, 3), "unit": "GB/s", "cores": threads, "host_cpus": host_cpus(),
            "kind": "port",
            "sample": "the oracle (scalar split-table C restatement) on the first %d bytes of all %d input slices x "
                      "%d recovery slices (%.3g mul-adds, %.1f s)" % (pwin, n, R, ma, t),
            "matches_gpu": bool((out == gpu_out[:, :pwin]).all())}
    if cpu is None:
        cpu = port
    else:
        cpu["port"] = port
    ref_here = reference_container_timing()
    if ref_here is not None and n == 2000 and S == 1 << 20 and R == 200:
        cpu["reference_in_build_container"] = {
            "value": ref_here["gbps_mul_add_equivalent"], "unit": "GB/s", "cores": ref_here["threads"],
            "method": ref_here["method"], "sample": "the whole workload, median of 3 runs (%s s)" % (
                ref_here["seconds"],), "host": ref_here["host"]}
    return cpu


def cpu_baseline_window(args, gpu_out):
    """Oracle on the first `window` bytes of every slice (all inputs, all recovery slices)."""
    import numpy as np
    import oracle as O

    n, R, S = args.inputs, args.recovery, args.slice
    exps = np.arange(R, dtype=np.uint16)
    threads = cpu_threads()

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
    win = max(tile, min(tiles * tile, S))
    win -= win % 1024
    log("cpu baseline: %d threads, window %d bytes of each slice" % (threads, win))
    t, out = run(win)
    ma = n * R * (win // 2)
    return {
        "value": round(2.0 * ma / t / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
        "sample": "first %d bytes of all %d input slices x %d recovery slices (%.3g mul-adds, %.1f s)" % (
            win, n, R, ma, t),
        "matches_gpu": bool((out == gpu_out[:, :win]).all()),
    }


def cpu_baseline_rows(args, gpu_rows, rows):
    """Oracle on a sample of recovery slices over the full inputs."""
    import numpy as np
    import oracle as O

    n, S = args.inputs, args.slice
    threads = cpu_threads()
    inp = np.empty((n, S), np.uint8)
    for i in range(n):
        inp[i] = O.prng_bytes(args.seed, 0, i, S)
    t0 = time.perf_counter()
    out = O.recover(inp, [S] * n, S, np.array(rows, np.uint16), threads=threads)
    t = time.perf_counter() - t0
    ma = n * len(rows) * (S // 2)
    return {
        "value": round(2.0 * ma / t / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
        "sample": "recovery slices %s over all %d input slices x %d B (%.3g mul-adds, %.1f s)" % (
            rows, n, S, ma, t),
        "matches_gpu": bool((out == gpu_rows).all()),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs on a single device (byte-range sharding across "
              "GPUs is a design note only); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.config in (0, 4):
        if not 0 <= args.hw_queues <= 32:
            sys.exit("--hw-queues must be in 0..32")
        if args.hw_queues:
            # read by the HIP runtime when it initialises, which nothing above has done yet
            os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.config == 0:
        return main_all(args)
    if args.config == 2:
        print(json.dumps(main_batched(args)), flush=True)
        return
    if args.config == 4:
        print(json.dumps(main_serving(args)), flush=True)
        return
    if args.form == "backend":
        return main_backend(args)
    if args.form == "hashes":
        return main_hashes(args)
    print(json.dumps(main_kernel(args)), flush=True)


def main_kernel(args):
    """One device-resident workload in one form (plain, dense or bilinear)."""
    import numpy as np
    import parpar_amd
    from parpar_amd import proc as P

    parpar_amd.require_device()
    L = parpar_amd.lib()
    chk = parpar_amd.check
    n, R, S, ck = args.inputs, args.recovery, args.slice, args.cksum
    dense = args.form in ("dense", "bilinear")
    ck_off, slot = P.slot_layout(S, ck)
    ma = n * R * (S // 2)
    log("workload: config %d, %s form, %d inputs x %d bytes, %d recovery slices, MA=%.4g" % (
        args.config, args.form, n, S, R, ma))

    name = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    L.pph_device_info(0, name, 256, ctypes.byref(cus))

    stream = ctypes.c_void_p()
    chk(L.pph_stream_create(ctypes.byref(stream)))
    src = P.DeviceBuffer(slot * n)      # data-form inputs (arrive in HBM)
    prepd = P.DeviceBuffer(slot * n)    # prepared inputs with checksums
    out = P.DeviceBuffer(slot * R)
    valid = P.DeviceBuffer(4 * R)
    d_logs = P.DeviceBuffer(2 * n)
    d_logs.upload(np.array([L.pph_input_log(i) for i in range(n)], np.uint16))
    d_exps = P.DeviceBuffer(2 * R)
    d_exps.upload(np.arange(R, dtype=np.uint16))
    chk(L.pph_fill_random(src.ptr, slot, n, S, args.seed, 0, 0, stream), "fill")
    chk(L.pph_stream_synchronize(stream))
    log("inputs generated in HBM")

    t_ker = ctypes.c_void_p()
    chk(L.pph_timer_create(ctypes.byref(t_ker)))

    def muladd(form, timed):
        if timed:
            chk(L.pph_timer_start(t_ker, stream))
        if form == "dense":
            chk(L.pph_muladd_dense(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, out.ptr, R, 0, stream), "dense")
        elif form == "bilinear":
            chk(L.pph_muladd_bilinear(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, out.ptr, R, 0, stream), "bilinear")
        else:
            chk(L.pph_muladd_plain(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, None, 0, out.ptr, R, 0,
                                   args.group, stream), "plain")
        if timed:
            chk(L.pph_timer_stop(t_ker, stream))

    def step(form, timed):
        chk(L.pph_prep_slots_from(prepd.ptr, src.ptr, slot, n, None, S, ck, 0 if form in ("dense", "bilinear") else 1,
                                  stream), "prep")
        muladd(form, timed)
        chk(L.pph_verify_slots(out.ptr, slot, R, S, ck, valid.ptr, stream), "verify")

    def run_steps(form, steps, warmup):
        for w in range(warmup):
            step(form, False)
            chk(L.pph_stream_synchronize(stream))
            log("%s warmup %d done" % (form, w + 1))
        ker = []
        chk(L.pph_synchronize())
        t0 = time.perf_counter()
        for s in range(steps):
            step(form, True)
            ms = ctypes.c_float()
            chk(L.pph_timer_elapsed_ms(t_ker, ctypes.byref(ms)))
            ker.append(ms.value)
            log("%s step %d: kernel %.2f ms" % (form, s + 1, ms.value))
        chk(L.pph_stream_synchronize(stream))
        wall = time.perf_counter() - t0
        ok = valid.download(4 * R).view(np.int32)
        if not (ok == 1).all():
            raise SystemExit("recovery checksum mismatch on the device (%s): %d of %d slices invalid" % (
                form, int((ok != 1).sum()), R))
        return wall * 1e3 / steps, sum(ker) / len(ker)

    if args.calib_copy:
        cbytes = args.calib_copy - args.calib_copy % 16
        ca = P.DeviceBuffer(cbytes)
        cb = P.DeviceBuffer(cbytes)
        chk(L.pph_copy(cb.ptr, ca.ptr, cbytes, stream), "copy")
        chk(L.pph_stream_synchronize(stream))
        log("calibration copy of %d bytes done" % cbytes)
        del ca, cb

    compare = None
    if dense and not args.no_compare:
        _, p_ker = run_steps("plain", 1, 0)
        compare = {"plain_kernel_ms": round(p_ker, 3), "plain_ma_per_s": round(ma / (p_ker / 1e3), 1)}
        p_rows = out.download(min(R, 16) * slot).reshape(-1, slot)[:, :S].copy()

    ms_per_step, ker_avg = run_steps(args.form, args.steps, args.warmup)
    value = 2.0 * ma / (ms_per_step / 1e3) / 1e9

    if dense:
        # MFMA ops per mul-add: 2 x 256 MACs (16x16 bit-matrix, int8) or 2 x 81 (bilinear
        # formula on the fp4 e2m1 MFMA kernel)
        per_ma = 512 if args.form == "dense" else 162
        fp4 = args.form == "bilinear"
        peak = FP4_PEAK_TFLOPS if fp4 else INT8_PEAK_TOPS
        achieved = per_ma * ma / (ker_avg / 1e3) / 1e12
        if args.form == "dense":
            kname = "k_muladd_dense (int8 MFMA)"
        else:
            kname = "pph_muladd_bilinear: k_blm_gemm (fp4 MFMA) + plane and recombination kernels"
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s" if fp4 else "TOP/s",
                "frac": round(achieved / peak, 4), "traffic": MEASURED_TRAFFIC.get((args.form, n, S, R)),
                "kernel": kname, "kernel_ms": round(ker_avg, 3), "algorithmic_mfma_ops": per_ma * ma,
                "operand_type": "fp4 e2m1" if fp4 else "int8"}
        if compare is not None:
            compare["dense_vs_plain"] = round(compare["plain_kernel_ms"] / ker_avg, 3)
            d_rows = out.download(min(R, 16) * slot).reshape(-1, slot)[:, :S]
            compare["dense_equals_plain_rows"] = bool((d_rows == p_rows).all())
    else:
        roof = plain_roofline(L, ma, ker_avg, MEASURED_TRAFFIC.get(("plain", n, S, R)), (n + R) * slot)

    cpu = None
    if not args.no_cpu:
        if args.config == 3:
            rows = [0, 1, 2, 3, 4, 5, 6, 7, 1000, 2047, 4095, 5000, 6000, 7000, 8190, 8191]
            rows = [r for r in rows if r < R]
            got = np.stack([out.download(S, off=r * slot) for r in rows])
            cpu = cpu_baseline_rows(args, got, rows)
        else:
            host = out.download(R * slot).reshape(R, slot)
            cpu = cpu_baseline_reference(args, host[:, :S])
        log("cpu baseline: %s" % cpu)

    res = {
        "metric": "GF16 mul-add GB/s (2 B per GF(2^16) mul-add), %s form" % args.form,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16 (GF(2^16) words)" + ((", %s 0/1 MFMA operands" % roof["operand_type"]) if dense else ""),
        "data": "synthetic (counter-based random bytes, seed %d)" % args.seed,
        "config": {"workload": "%s recovery: %d inputs x %d B slices, %d recovery slices (exp 0..%d)" % (
            args.form, n, S, R, R - 1), "config_index": args.config, "inputs": n, "slice_bytes": S,
            "recovery": R, "cksum_bytes": ck, "mul_adds": ma, "device": name.value.decode(), "cus": cus.value},
        "ma_per_s": round(ma / (ms_per_step / 1e3), 1),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if compare is not None:
        res["compare"] = compare
    L.pph_timer_destroy(t_ker)
    L.pph_stream_destroy(stream)
    if dense:
        L.pph_bilinear_release()
    return res


if __name__ == "__main__":
    main()
