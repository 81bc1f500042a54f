"""Serving engine (pph_serve_* in include/parpar_hip.h): jobs arrive over time
and run with continuous batching on one device.

A job is one file whose slices are PAR2 inputs.  For every job the engine
returns the per-slice MD5 / CRC-32 (the IFSC entries, hasher/hasher_base.h:154-176),
the file MD5 (hasher_base.h:178-186) and the recovery slices with their
gf16_cksum flags (what PAR2Proc produces for those inputs, gf16/controller.h:248-326).
"""
import ctypes

import numpy as np

from . import check, lib
from .proc import _ptr


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_compute", ctypes.c_uint), ("md5_quantum", ctypes.c_size_t),
                ("staging_bytes", ctypes.c_size_t), ("bilinear_workspace", ctypes.c_size_t),
                ("tuning", ctypes.c_void_p), ("md5_host_threads", ctypes.c_uint), ("upload_threads", ctypes.c_uint),
                ("md5_order", ctypes.c_uint)]


class ServeJobDesc(ctypes.Structure):
    _fields_ = [("file_size", ctypes.c_uint64), ("slice_size", ctypes.c_size_t), ("n_out", ctypes.c_uint),
                ("first_input", ctypes.c_uint), ("exps", ctypes.c_void_p), ("host_data", ctypes.c_void_p),
                ("seed", ctypes.c_uint64), ("job_id", ctypes.c_uint32), ("recovery_out", ctypes.c_void_p),
                ("valid", ctypes.c_void_p), ("slice_md5", ctypes.c_void_p), ("slice_crc", ctypes.c_void_p),
                ("file_md5", ctypes.c_void_p), ("ifsc", ctypes.c_void_p), ("packet_set_id", ctypes.c_void_p),
                ("packet_md5", ctypes.c_void_p)]


class ServeResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int), ("latency_s", ctypes.c_double),
                ("queue_s", ctypes.c_double), ("compute_s", ctypes.c_double), ("run_s", ctypes.c_double),
                ("mul_adds", ctypes.c_double), ("bytes", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("md5_rounds", ctypes.c_uint64),
                ("bilinear_jobs", ctypes.c_uint64), ("host_md5_jobs", ctypes.c_uint64), ("peak_jobs", ctypes.c_uint),
                ("peak_bytes", ctypes.c_uint64),
                ("arena_bytes", ctypes.c_uint64),
                ("mul_adds", ctypes.c_double), ("bytes", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FileJob:
    """One file and the host buffers its results land in.

    data: the file bytes (host) or None for device-generated inputs (slice k =
    the counter-based generator's (seed, job_id, k)).  keep_recovery=False
    leaves the recovery slices on the device (only their flags come back).
    keep_ifsc: also the slice hashes as IFSC entries (20 bytes per slice).
    packet_set_id (16 bytes): also each recovery packet's MD5.
    """

    def __init__(self, file_size, slice_size, n_out, exps=None, data=None, seed=0, job_id=0, first_input=0,
                 keep_recovery=True, keep_slice_hashes=True, keep_ifsc=False, packet_set_id=None):
        self.file_size = int(file_size)
        self.slice_size = int(slice_size)
        self.n_in = -(-self.file_size // self.slice_size)
        self.n_out = int(n_out)
        self.exps = None if exps is None else np.ascontiguousarray(exps, np.uint16)
        self.data = None if data is None else np.ascontiguousarray(data, np.uint8)
        if self.data is not None:
            assert self.data.nbytes == self.file_size
        self.seed, self.job_id, self.first_input = int(seed), int(job_id), int(first_input)
        self.recovery = np.zeros((self.n_out, self.slice_size), np.uint8) if keep_recovery and self.n_out else None
        self.valid = np.full(max(self.n_out, 1), -1, np.int32)
        self.slice_md5 = np.zeros((self.n_in, 16), np.uint8) if keep_slice_hashes else None
        self.slice_crc = np.zeros(self.n_in, np.uint32) if keep_slice_hashes else None
        self.file_md5 = np.zeros(16, np.uint8)
        self.ifsc = np.zeros((self.n_in, 20), np.uint8) if keep_ifsc else None
        self.packet_set_id = None if packet_set_id is None else np.frombuffer(bytes(packet_set_id), np.uint8).copy()
        if self.packet_set_id is not None:
            assert self.packet_set_id.size == 16
        self.packet_md5 = np.zeros((max(self.n_out, 1), 16), np.uint8) if packet_set_id is not None else None
        self.result = None

    def desc(self):
        p = lambda a: None if a is None else _ptr(a)
        return ServeJobDesc(self.file_size, self.slice_size, self.n_out, self.first_input, p(self.exps),
                            p(self.data), self.seed, self.job_id, p(self.recovery), p(self.valid),
                            p(self.slice_md5), p(self.slice_crc), p(self.file_md5), p(self.ifsc),
                            p(self.packet_set_id), p(self.packet_md5))

    @property
    def mul_adds(self):
        return self.n_in * self.n_out * (self.slice_size // 2)


class Server:
    """A pph_serve engine.  Jobs (FileJob) are kept alive until their result
    has been polled, so their host buffers stay valid while the device writes
    them."""

    def __init__(self, hbm_budget, max_compute=8, md5_quantum=0, staging_bytes=0, device=0, bilinear_workspace=0,
                 tuning=None, md5_host_threads=0, upload_threads=0, md5_order=0):
        """bilinear_workspace > 0 lets large jobs use the bilinear form; tuning
        (parpar_amd.batch.PPHTuning) gives the rates that choose it, None
        measures them at create.  md5_host_threads > 0 computes every file MD5
        on that many host threads (device-generated files read back from
        HBM); upload_threads copy host bytes to the device (0 = 1); md5_order
        0 hashes the queued file with the highest response ratio first, 1 in
        submit order."""
        self._tuning = tuning  # kept alive while the engine reads it
        self.cfg = ServeConfig(int(hbm_budget), int(max_compute), int(md5_quantum), int(staging_bytes),
                               int(bilinear_workspace),
                               ctypes.cast(ctypes.pointer(tuning), ctypes.c_void_p) if tuning is not None else None,
                               int(md5_host_threads), int(upload_threads), int(md5_order))
        h = ctypes.c_void_p()
        check(lib().pph_serve_create(device, ctypes.byref(self.cfg), ctypes.byref(h)), "pph_serve_create")
        self.h = h
        self.jobs = {}

    def submit(self, job):
        t = ctypes.c_uint64()
        d = job.desc()
        check(lib().pph_serve_submit(self.h, ctypes.byref(d), ctypes.byref(t)), "pph_serve_submit")
        self.jobs[t.value] = job
        return t.value

    def poll(self, wait_ms=0, max_results=256):
        """Finished jobs since the last poll, each with .result set."""
        buf = (ServeResult * max_results)()
        n = ctypes.c_uint()
        check(lib().pph_serve_poll(self.h, buf, max_results, ctypes.byref(n), int(wait_ms)), "pph_serve_poll")
        out = []
        for k in range(n.value):
            r = buf[k]
            job = self.jobs.pop(r.ticket)
            job.result = r.as_dict()
            out.append(job)
        return out

    def drain(self, wait_ms):
        return check(lib().pph_serve_drain(self.h, int(wait_ms)), "pph_serve_drain")

    def stats(self):
        s = ServeStats()
        check(lib().pph_serve_get_stats(self.h, ctypes.byref(s)), "pph_serve_get_stats")
        return s.as_dict()

    def close(self):
        if self.h is not None:
            lib().pph_serve_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run_all(self, jobs, timeout_s=120):
        """Submit every job, wait for all of them; returns them in order."""
        for j in jobs:
            self.submit(j)
        left = self.drain(int(timeout_s * 1000))
        if left:
            raise RuntimeError("%d serving jobs did not finish within %.0f s" % (left, timeout_s))
        done = self.poll(max_results=max(len(jobs), 1))
        assert len(done) == len(jobs)
        return jobs
