"""Batched execution of many PAR2 jobs (pph_batch_* in include/parpar_hip.h).

A job is device-resident: input slots (data form, prepared in place by every
run), recovery slots, per-input lengths, PAR2 input logs, recovery exponents
and checksum flags.  ``synthetic_job`` builds one from the counter-based input
generator, the way the benchmark's batched workload does.
"""
import ctypes

import numpy as np

from . import check, lib
from .proc import DeviceBuffer, slot_layout


class PPHJob(ctypes.Structure):
    _fields_ = [("in_slots", ctypes.c_void_p), ("out_slots", ctypes.c_void_p), ("in_lens", ctypes.c_void_p),
                ("in_logs", ctypes.c_void_p), ("exps", ctypes.c_void_p), ("valid", ctypes.c_void_p),
                ("slice_size", ctypes.c_size_t), ("slot_bytes", ctypes.c_size_t), ("n_in", ctypes.c_uint),
                ("n_out", ctypes.c_uint)]


class PPHTuning(ctypes.Structure):
    _fields_ = [("version", ctypes.c_uint), ("r_buckets", ctypes.c_uint), ("plain_g8", ctypes.c_float * 8),
                ("plain_g16", ctypes.c_float * 8), ("dense", ctypes.c_float * 8), ("bilinear", ctypes.c_float * 8),
                ("bilinear_wide", ctypes.c_float * 8)]

    def as_dict(self):
        return {"plain_g8": list(self.plain_g8), "plain_g16": list(self.plain_g16), "dense": list(self.dense),
                "bilinear": list(self.bilinear), "bilinear_wide": list(self.bilinear_wide)}


def autotune():
    t = PPHTuning()
    check(lib().pph_autotune(ctypes.byref(t)), "pph_autotune")
    return t


def fixed_tuning(plain=1.0, dense=0.0, bilinear=0.0, g16_over_g8=1.01):
    """A tuning table that makes the planner pick a given form (tests)."""
    t = PPHTuning()
    t.version, t.r_buckets = 3, 8
    for b in range(8):
        t.plain_g8[b] = plain * 1e12
        t.plain_g16[b] = plain * 1e12 * g16_over_g8
        t.dense[b] = dense * 1e12
        t.bilinear[b] = bilinear * 1e12
        t.bilinear_wide[b] = bilinear * 1e12
    return t


class Job:
    """One device-resident PAR2 job."""

    def __init__(self, n_in, slice_size, exps, input_nums=None, lens=None, cksum=64):
        self.n_in = int(n_in)
        self.slice_size = int(slice_size)
        self.exps = np.asarray(exps, np.uint16)
        self.n_out = len(self.exps)
        self.input_nums = np.arange(self.n_in, dtype=np.uint16) if input_nums is None else \
            np.asarray(input_nums, np.uint16)
        self.lens = np.full(self.n_in, self.slice_size, np.uint32) if lens is None else np.asarray(lens, np.uint32)
        self.ck_off, self.slot = slot_layout(self.slice_size, cksum)
        self.d_in = DeviceBuffer(self.slot * max(self.n_in, 1))
        self.d_out = DeviceBuffer(self.slot * max(self.n_out, 1))
        self.d_lens = DeviceBuffer(4 * max(self.n_in, 1))
        self.d_lens.upload(self.lens)
        self.d_logs = DeviceBuffer(2 * max(self.n_in, 1))
        self.d_logs.upload(np.array([lib().pph_input_log(int(x)) for x in self.input_nums], np.uint16))
        self.d_exps = DeviceBuffer(2 * max(self.n_out, 1))
        self.d_exps.upload(self.exps)
        self.d_valid = DeviceBuffer(4 * max(self.n_out, 1))

    def desc(self):
        return PPHJob(self.d_in.addr, self.d_out.addr, self.d_lens.addr, self.d_logs.addr, self.d_exps.addr,
                      self.d_valid.addr, self.slice_size, self.slot, self.n_in, self.n_out)

    def fill(self, seed, job_id, stream=None):
        """Inputs from the counter-based generator: slice k of job `job_id`."""
        check(lib().pph_fill_random(self.d_in.ptr, self.slot, self.n_in, self.slice_size, seed, job_id, 0, stream),
              "pph_fill_random")

    def outputs(self):
        return self.d_out.download(self.slot * self.n_out).reshape(self.n_out, self.slot)[:, :self.slice_size]

    def valid(self):
        return self.d_valid.download(4 * self.n_out).view(np.int32).copy()


class Batch:
    def __init__(self, jobs, cksum=64, tuning=None):
        self.jobs = list(jobs)
        arr = (PPHJob * max(len(self.jobs), 1))(*[j.desc() for j in self.jobs])
        p = ctypes.c_void_p()
        check(lib().pph_batch_create(ctypes.cast(arr, ctypes.c_void_p), len(self.jobs), cksum,
                                     ctypes.byref(tuning) if tuning is not None else None, ctypes.byref(p)),
              "pph_batch_create")
        self._p = p

    def plan(self, j):
        form = ctypes.c_int()
        group = ctypes.c_uint()
        pred = ctypes.c_double()
        check(lib().pph_batch_plan(self._p, j, ctypes.byref(form), ctypes.byref(group), ctypes.byref(pred)),
              "pph_batch_plan")
        return {"form": {1: "plain", 2: "dense", 3: "bilinear"}.get(form.value, form.value), "group": group.value,
                "predicted_s": pred.value}

    def run(self, use_graph=True, stream=None):
        check(lib().pph_batch_run(self._p, 1 if use_graph else 0, stream), "pph_batch_run")

    def fused_ms(self):
        """Mul-add time of the last eager run (HIP events on its stream)."""
        ms = ctypes.c_float()
        check(lib().pph_batch_fused_ms(self._p, ctypes.byref(ms)), "pph_batch_fused_ms")
        return ms.value

    def close(self):
        if self._p is not None:
            lib().pph_batch_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
