"""The entry point without allocations, pph_muladd_bilinear_noalloc, through
the C ABI: its recovery slots (data and checksum blocks) equal
pph_muladd_bilinear's byte for byte on calls in a row, each smaller than the
one before, overwriting and accumulating, up to 32768 inputs (the static
scratch then holds one 128-output by 512-word chunk at a time); the device's
free memory is the same before and after its calls."""
import ctypes

import numpy as np
import pytest

import parpar_amd
from parpar_amd import proc as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def lib():
    return parpar_amd.lib()


def free_bytes():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    parpar_amd.check(lib().pph_mem_get_info(ctypes.byref(f), ctypes.byref(t)), "mem_get_info")
    return f.value


class Case:
    def __init__(self, n, n_out, S, seed):
        L = lib()
        self.n, self.n_out, self.S = n, n_out, S
        _, self.slot = P.slot_layout(S, 64)
        src = P.DeviceBuffer(self.slot * n)
        parpar_amd.check(L.pph_fill_random(src.ptr, self.slot, n, S, seed, 0, 0, None), "fill")
        self.data = P.DeviceBuffer(self.slot * n)
        parpar_amd.check(L.pph_prep_slots_from(self.data.ptr, src.ptr, self.slot, n, None, S, 64, 0, None), "prep")
        src.close()
        rng = np.random.default_rng(seed)
        nums = rng.permutation(32768)[:n].astype(np.uint16)
        self.logs = P.DeviceBuffer(2 * n)
        self.logs.upload(np.array([L.pph_input_log(int(x)) for x in nums], np.uint16))
        self.exps = P.DeviceBuffer(2 * n_out)
        self.exps.upload(rng.integers(0, 65535, n_out).astype(np.uint16))
        self.start = np.frombuffer(rng.bytes(self.slot * n_out), np.uint8)
        self.ref = P.DeviceBuffer(self.slot * n_out)
        self.got = P.DeviceBuffer(self.slot * n_out)
        parpar_amd.check(L.pph_synchronize())

    def run(self, fn, out, accumulate):
        out.upload(self.start)
        parpar_amd.check(fn(self.data.ptr, self.slot, self.n, self.logs.ptr, self.exps.ptr, out.ptr, self.n_out,
                            accumulate, None), fn.__name__)


def test_noalloc_equals_bilinear_on_calls_in_a_row():
    L = lib()
    # each call smaller than the one before it
    cases = [Case(32768, 300, 1000, 11), Case(1100, 260, 20000, 12), Case(40, 17, 5000, 13)]
    first = True
    for c in cases:
        for acc in (0, 1):
            say("case n=%d n_out=%d S=%d accumulate=%d" % (c.n, c.n_out, c.S, acc))
            c.run(L.pph_muladd_bilinear, c.ref, acc)
            parpar_amd.check(L.pph_synchronize())
            before = free_bytes()
            c.run(L.pph_muladd_bilinear_noalloc, c.got, acc)
            parpar_amd.check(L.pph_synchronize())
            delta = before - free_bytes()
            say("  free device memory change: %d bytes%s" % (delta, " (first call: set-up)" if first else ""))
            # the first call on the device is the set-up (symbol addresses, exp table)
            assert first or delta == 0
            first = False
            want = c.ref.download().reshape(c.n_out, c.slot)
            got = c.got.download().reshape(c.n_out, c.slot)
            bad = np.flatnonzero((want != got).any(axis=1))
            assert bad.size == 0, (c.n, c.n_out, c.S, acc, bad[:8])
    L.pph_bilinear_release()


def test_noalloc_rejects_bad_shapes():
    L = lib()
    buf = P.DeviceBuffer(4096)
    assert L.pph_muladd_bilinear_noalloc(buf.ptr, 1000, 1, buf.ptr, buf.ptr, buf.ptr, 1, 0, None) == \
        parpar_amd.PPH_E_INVALID
    assert L.pph_muladd_bilinear_noalloc(buf.ptr, 1024, 32769, buf.ptr, buf.ptr, buf.ptr, 1, 0, None) == \
        parpar_amd.PPH_E_INVALID
    assert L.pph_muladd_bilinear_noalloc(buf.ptr, 1024, 1, None, None, buf.ptr, 1, 0, None) == \
        parpar_amd.PPH_E_INVALID
