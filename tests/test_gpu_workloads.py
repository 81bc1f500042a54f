"""Parity on the batched and serving workloads at their full sizes.

Batched: the 64 uneven jobs of seed 2 (about 84 GiB of inputs on the device)
planned by the autotuner and run from the batch's hipGraph, replayed twice.
After each replay every recovery slice's checksum flag is valid, every
checksum block equals the oracle's recovery over the inputs' checksum blocks
(the checksum is linear, so this pins every row of every job), and a 4 KiB
column window of every job equals the oracle's recovery of that window.

Serving: the first 3 s of the seed-4 Poisson stream (41 files, 42 GiB),
arriving on time, with every file MD5 on the device (shared hash rounds, as
the serving workload specifies).  Every file MD5, slice MD5 and slice CRC-32
equals hashlib / zlib over the generator's bytes, every recovery flag is
valid, and a 4 KiB column window of every file's recovery equals the
oracle's."""
import hashlib
import threading
import time
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import batch as B
from parpar_amd import serve as SV
from parpar_amd import workloads as W

pytestmark = pytest.mark.gpu

THREADS = 16  # the GPU host's CPU share


def say(*a):
    print(*a, flush=True)


def window_check(seed, job_id, n, S, lens, exps, off, win, got):
    """The oracle's recovery of bytes [off, off + win) of every input against got."""
    inp = np.zeros((n, win), np.uint8)
    wl = [max(0, min(win, int(l) - off)) for l in lens]
    for k in range(n):
        if wl[k]:
            inp[k, :wl[k]] = O.prng_bytes(seed, job_id, k, wl[k], offset=off)
    ref = O.recover(inp, wl, win, exps, threads=THREADS)
    return bool((ref == got).all())


def test_batched_64_jobs_graph_replays():
    shapes = W.batched_jobs(2)
    jobs = []
    for n, S, last, R in shapes:
        lens = [S] * n
        lens[-1] = last
        jobs.append(B.Job(n, S, np.arange(R), lens=lens))
    say("batched: %d jobs, %.1f GiB of input slots" % (len(jobs), sum(j.slot * j.n_in for j in jobs) / 2**30))
    tuning = B.autotune()
    batch = B.Batch(jobs, tuning=tuning)
    forms = [batch.plan(j)["form"] for j in range(len(jobs))]
    say("plan:", {f: forms.count(f) for f in set(forms)})
    # the oracle's checksum blocks of every job's recovery (inputs' blocks through the recovery)
    want_ck = []
    for j, job in enumerate(jobs):
        cks = O.prng_cksums(2, j, job.n_in, job.slice_size, lens=job.lens, threads=THREADS)
        want_ck.append(O.recover(cks, [64] * job.n_in, 64, job.exps, threads=THREADS))
    say("oracle checksum blocks ready")
    rng = np.random.default_rng(77)
    for replay in range(2):
        for j, job in enumerate(jobs):
            job.fill(2, j)
        parpar_amd.check(parpar_amd.lib().pph_synchronize())
        batch.run(use_graph=True)
        parpar_amd.check(parpar_amd.lib().pph_synchronize())
        say("replay %d done" % replay)
        for j, job in enumerate(jobs):
            v = job.valid()
            assert (v == 1).all(), (replay, j, np.flatnonzero(v != 1)[:8])
            got_ck = np.stack([job.d_out.download(64, off=r * job.slot + job.ck_off) for r in range(job.n_out)])
            bad = np.flatnonzero((got_ck != want_ck[j]).any(axis=1))
            assert bad.size == 0, (replay, j, forms[j], bad[:8])
            win = min(4096, job.slice_size)
            off = int(rng.integers(0, job.slice_size - win + 1)) & ~1
            got = np.stack([job.d_out.download(win, off=r * job.slot + off) for r in range(job.n_out)])
            assert window_check(2, j, job.n_in, job.slice_size, job.lens, job.exps, off, win, got), (replay, j)
        say("replay %d: every flag, checksum block and window matches" % replay)
    batch.close()
    parpar_amd.lib().pph_bilinear_release()


def file_hashes(seed, job_id, file_size, S, n):
    """File MD5, slice MD5s and CRC-32s over the generator's bytes (hashlib / zlib)."""
    fm = hashlib.md5()
    md5s, crcs = [], []
    for k in range(n):
        L = min(S, file_size - k * S)
        d = O.prng_bytes(seed, job_id, k, L)
        fm.update(d)
        pad = d.tobytes() + bytes(S - L)
        md5s.append(hashlib.md5(pad).digest())
        crcs.append(zlib.crc32(pad) & 0xFFFFFFFF)
    return fm.digest(), md5s, crcs


def test_serving_3s_window_device_file_md5():
    arrivals = W.serving_arrivals(4, 3.0)
    L = parpar_amd.lib()
    import ctypes
    free, tot = ctypes.c_size_t(), ctypes.c_size_t()
    parpar_amd.check(L.pph_mem_get_info(ctypes.byref(free), ctypes.byref(tot)))
    bl_ws = 40 << 30
    budget = min(free.value - (16 << 30) - bl_ws, 200 << 30)
    say("serving: %d files, %.1f GiB; arena %.0f GiB; file MD5 on the device" % (
        len(arrivals), sum(a[1] for a in arrivals) / 2**30, budget / 2**30))
    srv = SV.Server(budget, max_compute=8, bilinear_workspace=bl_ws, md5_host_threads=0)
    jobs = [SV.FileJob(T, S, R, seed=4, job_id=k) for k, (_, T, S, R) in enumerate(arrivals)]
    t0 = time.perf_counter()

    def submitter():
        for k, a in enumerate(arrivals):
            d = t0 + a[0] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            srv.submit(jobs[k])

    th = threading.Thread(target=submitter)
    th.start()
    # the CPU side of the check runs while the device works
    pool = ThreadPoolExecutor(THREADS)
    futs = {k: pool.submit(file_hashes, 4, k, T, S, -(-T // S)) for k, (_, T, S, _) in enumerate(arrivals)}
    done = 0
    last = t0
    while done < len(arrivals):
        for job in srv.poll(wait_ms=1000):
            assert job.result["status"] == 0, job.result
            done += 1
        if time.perf_counter() - last > 10:
            last = time.perf_counter()
            say("t=%.0f s: %d/%d files done" % (last - t0, done, len(arrivals)))
    th.join()
    say("all files done in %.1f s; stats %s" % (time.perf_counter() - t0, srv.stats()))
    srv.close()
    rng = np.random.default_rng(78)
    for k, (_, T, S, R) in enumerate(arrivals):
        job = jobs[k]
        fmd5, smd5, scrc = futs[k].result()
        assert bytes(job.file_md5) == fmd5, k
        for i in range(job.n_in):
            assert bytes(job.slice_md5[i]) == smd5[i], (k, i)
            assert int(job.slice_crc[i]) == scrc[i], (k, i)
        assert (job.valid[:R] == 1).all(), k
        win = min(4096, S)
        off = int(rng.integers(0, S - win + 1)) & ~1
        lens = [S] * job.n_in
        lens[-1] = T - (job.n_in - 1) * S
        assert window_check(4, k, job.n_in, S, lens, np.arange(R, dtype=np.uint16), off, win,
                            job.recovery[:, off:off + win]), k
    pool.shutdown()
    say("every file's hashes, flags and recovery window match")
