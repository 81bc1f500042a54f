"""Serving engine: jobs submitted over time, admitted by device-memory
footprint, run concurrently; every output (recovery slices, checksum flags,
per-slice MD5 / CRC-32, file MD5) checked against the oracle and
hashlib / zlib, and a job's outputs shown to be the same alone and mixed with
other jobs.  Calls go through the C ABI (pph_serve_*)."""
import ctypes
import hashlib
import zlib

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import batch as B
from parpar_amd import serve as SV
from golden_io import expected_valid

pytestmark = pytest.mark.gpu
MiB = 1 << 20


def say(*a):
    print(*a, flush=True)


def file_bytes(job):
    if job.data is not None:
        return job.data
    S = job.slice_size
    parts = [O.prng_bytes(job.seed, job.job_id, k, min(S, job.file_size - k * S)) for k in range(job.n_in)]
    return np.concatenate(parts)


def check_job(job, data=None):
    """Compare a finished job with the oracle and Python's hash routines."""
    assert job.result is not None and job.result["status"] == 0, job.result
    data = file_bytes(job) if data is None else data
    S, n = job.slice_size, job.n_in
    assert bytes(job.file_md5).hex() == hashlib.md5(data.tobytes()).hexdigest()
    padded = np.zeros((n, S), np.uint8)
    padded.reshape(-1)[:job.file_size] = data
    lens = np.full(n, S, np.uint32)
    lens[-1] = job.file_size - (n - 1) * S
    if job.slice_md5 is not None:
        for k in range(n):
            row = padded[k].tobytes()
            assert bytes(job.slice_md5[k]).hex() == hashlib.md5(row).hexdigest(), k
            assert int(job.slice_crc[k]) == zlib.crc32(row) & 0xffffffff, k
    if job.n_out:
        exps = np.arange(job.n_out, dtype=np.uint16) if job.exps is None else job.exps
        nums = np.arange(job.first_input, job.first_input + n, dtype=np.uint16)
        if job.recovery is not None:
            ref = O.recover(padded, lens, S, exps, input_nums=nums)
            assert (job.recovery == ref).all()
        flags = expected_valid(padded, lens, S, exps, nums)
        assert ((job.valid[:job.n_out] == 1) == flags).all()


def mixed_jobs():
    """A spread of shapes: host and device inputs, partial and odd last
    slices, a single slice, hashes only, explicit exponents, input offsets."""
    rng = np.random.default_rng(44)
    jobs = []
    specs = [
        # file_size, slice_size, n_out, host?, exps, first_input
        (5 * 65536 + 1001, 65536, 3, True, None, 0),
        (40 * 4096, 4096, 7, False, None, 0),
        (999, 1024, 2, True, [0, 65534], 0),
        (3, 4, 1, False, None, 0),
        (17 * 8192 - 7, 8192, 5, False, [1, 2, 3, 100, 200], 100),
        (100000, 20000, 0, True, None, 0),
        (64 * 1024 * 3, 64 * 1024, 12, False, list(range(12)), 32760),
        (2 * MiB + 12, MiB, 2, True, None, 7),
    ]
    for j, (fs, S, R, host, exps, first) in enumerate(specs):
        data = O.prng_bytes(7, 1000 + j, 0, fs) if host else None
        jobs.append(SV.FileJob(fs, S, R, exps=exps, data=data, seed=5, job_id=j, first_input=first))
    rng.shuffle(jobs)
    return jobs


def test_serve_mixed_jobs():
    jobs = mixed_jobs()
    say("serve mixed", [(j.file_size, j.slice_size, j.n_out) for j in jobs])
    with SV.Server(hbm_budget=256 * MiB, max_compute=4, md5_quantum=64 * 1024) as srv:
        srv.run_all(jobs)
        st = srv.stats()
    say(" stats", st)
    assert st["completed"] == len(jobs) and st["peak_bytes"] <= 256 * MiB
    for job in jobs:
        check_job(job)


def test_serve_ifsc_and_packet_md5():
    """IFSC entries and recovery-packet MD5s from the engine: against hashlib /
    zlib over the padded slices and over the recovery packets assembled as
    lib/par2.js:505-513 writes them, for host and device inputs, both forms."""
    import struct
    set_id = bytes(range(100, 116))
    say("serve ifsc + packet md5")
    for tuning in (None, B.fixed_tuning(plain=0.001, dense=0.0, bilinear=1.0)):
        jobs = []
        for j, (fs, S, R, host) in enumerate([(5 * 65536 + 1001, 65536, 3, True), (40 * 4096, 4096, 7, False),
                                               (999, 1024, 2, True), (17 * 8192 - 6, 8192, 5, False)]):
            data = O.prng_bytes(7, 2000 + j, 0, fs) if host else None
            jobs.append(SV.FileJob(fs, S, R, data=data, seed=6, job_id=j, first_input=j, keep_ifsc=True,
                                   packet_set_id=set_id))
        kw = {} if tuning is None else dict(bilinear_workspace=4 << 30, tuning=tuning)
        with SV.Server(hbm_budget=128 * MiB, max_compute=3, **kw) as srv:
            srv.run_all(jobs)
        for job in jobs:
            check_job(job)
            data = file_bytes(job)
            S = job.slice_size
            padded = np.zeros(job.n_in * S, np.uint8)
            padded[:job.file_size] = data
            for k in range(job.n_in):
                row = padded[k * S:(k + 1) * S].tobytes()
                assert bytes(job.ifsc[k]) == hashlib.md5(row).digest() + struct.pack("<I", zlib.crc32(row)), k
            for r in range(job.n_out):
                pkt = set_id + b"PAR 2.0\0RecvSlic" + struct.pack("<I", r) + job.recovery[r].tobytes()
                assert bytes(job.packet_md5[r]) == hashlib.md5(pkt).digest(), r


def test_serve_alone_vs_mixed():
    """A job's outputs do not depend on what runs beside it."""
    def target():
        return SV.FileJob(9 * 32768 + 123, 32768, 6, seed=11, job_id=77, first_input=3)

    alone = target()
    with SV.Server(hbm_budget=64 * MiB, max_compute=1) as srv:
        srv.run_all([alone])
    check_job(alone)

    mixed = target()
    others = mixed_jobs()
    # the target sits among the others; a tight budget forces queueing and bypass
    order = others[:3] + [mixed] + others[3:]
    with SV.Server(hbm_budget=24 * MiB, max_compute=3, md5_quantum=4096) as srv:
        srv.run_all(order)
        st = srv.stats()
    say(" mixed stats", st)
    for name in ("recovery", "valid", "slice_md5", "slice_crc", "file_md5"):
        assert (getattr(alone, name) == getattr(mixed, name)).all(), name
    for job in others:
        check_job(job)


def test_serve_bilinear_form():
    """Jobs routed to the bilinear form (forced by the rate table) give the
    oracle's bytes, alone and mixed, and the same bytes as the plain form."""
    say("serve bilinear")
    force = B.fixed_tuning(plain=0.001, dense=0.0, bilinear=1.0)
    jobs = mixed_jobs()
    with SV.Server(hbm_budget=256 * MiB, max_compute=4, md5_quantum=64 * 1024, bilinear_workspace=12 << 30,
                   tuning=force) as srv:
        srv.run_all(jobs)
        st = srv.stats()
    say(" stats", st)
    assert st["bilinear_jobs"] == sum(1 for j in jobs if j.n_out)
    for job in jobs:
        check_job(job)

    def target():
        return SV.FileJob(9 * 32768 + 123, 32768, 6, seed=11, job_id=77, first_input=3)

    plain = target()
    with SV.Server(hbm_budget=64 * MiB, max_compute=1) as srv:
        srv.run_all([plain])
    bl = target()
    with SV.Server(hbm_budget=64 * MiB, max_compute=2, bilinear_workspace=4 << 30, tuning=force) as srv:
        srv.run_all(mixed_jobs()[:4] + [bl])
        assert srv.stats()["bilinear_jobs"] >= 1
    for name in ("recovery", "valid", "slice_md5", "slice_crc", "file_md5"):
        assert (getattr(plain, name) == getattr(bl, name)).all(), name


def test_serve_host_md5_and_uploaders():
    """File MD5 on host threads (host-sourced files from host memory,
    device-generated files read back from HBM in small pinned chunks, several
    pieces per slice) and host bytes copied by several upload threads give the
    same bytes as device hashing, alone and mixed."""
    jobs = mixed_jobs()
    say("serve host md5", len(jobs))
    with SV.Server(hbm_budget=256 * MiB, max_compute=4, staging_bytes=12288, md5_host_threads=3,
                   upload_threads=3) as srv:
        srv.run_all(jobs)
        st = srv.stats()
    say(" stats", st)
    assert st["host_md5_jobs"] == len(jobs) and st["md5_rounds"] == 0
    for job in jobs:
        check_job(job)

    # device-generated file, hashed on the device and on a host thread
    def generated():
        return SV.FileJob(13 * 20000 + 44, 20000, 3, seed=21, job_id=9, first_input=5)

    on_dev, on_host = generated(), generated()
    with SV.Server(hbm_budget=64 * MiB, max_compute=2, md5_quantum=4096) as srv:
        srv.run_all([on_dev])
    with SV.Server(hbm_budget=64 * MiB, max_compute=2, md5_host_threads=2, staging_bytes=7000) as srv:
        srv.run_all([on_host] + mixed_jobs()[:3])
        assert srv.stats()["md5_rounds"] == 0
    for name in ("recovery", "valid", "slice_md5", "slice_crc", "file_md5"):
        assert (getattr(on_dev, name) == getattr(on_host, name)).all(), name
    check_job(on_host)

    def target():
        fs = 11 * 16384 + 5
        return SV.FileJob(fs, 16384, 4, data=O.prng_bytes(9, 3, 0, fs), first_input=40)

    dev = target()
    with SV.Server(hbm_budget=64 * MiB, max_compute=1, md5_quantum=4096) as srv:
        srv.run_all([dev])
        assert srv.stats()["host_md5_jobs"] == 0
    host = target()
    with SV.Server(hbm_budget=24 * MiB, max_compute=3, md5_host_threads=1, upload_threads=2,
                   staging_bytes=5000) as srv:
        srv.run_all(mixed_jobs()[:5] + [host] + mixed_jobs()[5:])
    for name in ("recovery", "valid", "slice_md5", "slice_crc", "file_md5"):
        assert (getattr(dev, name) == getattr(host, name)).all(), name
    check_job(host)


@pytest.mark.parametrize("order", [0, 1])
def test_serve_host_md5_order(order):
    """One host MD5 thread, busy with a first file while a large and then a
    small file queue: by response ratio (md5_order 0) the small one is hashed
    first, in submit order (1) after the large one.  Every output checks."""
    say("host md5 order", order)
    first, large, small = (SV.FileJob(fs, MiB, 0, data=O.prng_bytes(12, k, 0, fs), job_id=k)
                           for k, fs in enumerate([160 * MiB, 320 * MiB, MiB + 100]))
    done = []
    with SV.Server(hbm_budget=2048 * MiB, max_compute=4, md5_host_threads=1, upload_threads=2,
                   md5_order=order) as srv:
        for job in (first, large, small):
            srv.submit(job)
        while len(done) < 3:
            done += [j.job_id for j in srv.poll(wait_ms=1000)]
    say(" completion order", done)
    assert done.index(small.job_id) < done.index(large.job_id) if order == 0 else \
        done.index(small.job_id) > done.index(large.job_id)
    for job in (first, large, small):
        check_job(job)


@pytest.mark.parametrize("quantum", [64, 4096, 8 * MiB])
def test_serve_md5_rounds(quantum):
    """Files of many sizes hash in shared rounds of `quantum` bytes per chain."""
    sizes = [1, 63, 64, 65, 4095, 4096, 4097, 100003, 3 * 65536]
    jobs = [SV.FileJob(fs, 65536, 0, data=O.prng_bytes(8, k, 0, fs), job_id=k) for k, fs in enumerate(sizes)]
    say("md5 rounds", quantum)
    with SV.Server(hbm_budget=64 * MiB, max_compute=8, md5_quantum=quantum) as srv:
        srv.run_all(jobs)
        rounds = srv.stats()["md5_rounds"]
    say(" rounds", rounds)
    for job in jobs:
        assert bytes(job.file_md5).hex() == hashlib.md5(job.data.tobytes()).hexdigest(), job.file_size
    if quantum == 64:
        assert rounds >= 3 * 65536 // 64


def test_serve_rejects():
    with SV.Server(hbm_budget=8 * MiB, max_compute=2) as srv:
        lib = parpar_amd.lib()
        t = ctypes.c_uint64()

        def submit(job):
            d = job.desc()
            return lib.pph_serve_submit(srv.h, ctypes.byref(d), ctypes.byref(t))

        assert submit(SV.FileJob(16 * MiB, MiB, 1)) == parpar_amd.PPH_E_NOMEM  # never fits
        assert submit(SV.FileJob(1000, 6, 1)) == parpar_amd.PPH_E_INVALID      # slice not a multiple of 4
        assert submit(SV.FileJob(4 * 32769, 4, 1)) == parpar_amd.PPH_E_INVALID  # more than 32768 slices
        assert submit(SV.FileJob(400, 4, 1, first_input=32700)) == parpar_amd.PPH_E_INVALID
        assert submit(SV.FileJob(400, 4, 1, exps=[65535])) == parpar_amd.PPH_E_INVALID
        ok = SV.FileJob(4000, 1024, 2, seed=1, job_id=1)
        srv.run_all([ok])
    check_job(ok)
