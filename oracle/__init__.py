"""ORACLE — test infrastructure, not product code.

ctypes bindings for the C restatement (``oracle/liboracle.so``) and, where it
has been built in this container, the reference harness
(``oracle/_ref/libparpar_ref.so``, compiled from the reference's own gf16/
sources).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline leg import this package, and only to check or time the CPU side.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libparpar_ref.so")
# the same harness over a portable build of the reference (per-file ISA flags, run-time
# method choice by CPUID): the one to run on a host other than the build container
REF_PORTABLE_SO = os.path.join(_HERE, "_ref", "libparpar_ref_portable.so")

_c_u8p = ctypes.POINTER(ctypes.c_uint8)
_c_u16p = ctypes.POINTER(ctypes.c_uint16)
_c_szp = ctypes.POINTER(ctypes.c_size_t)
_c_intp = ctypes.POINTER(ctypes.c_int)

_lib = None
_ref = None


def _ptr(a, t=_c_u8p):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def lib():
    """Load the C restatement (building it is ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(ORACLE_SO)
        L.po_mul_slow.restype = ctypes.c_uint16
        L.po_mul_slow.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.po_mul.restype = ctypes.c_uint16
        L.po_mul.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.po_exp.restype = ctypes.c_uint16
        L.po_exp.argtypes = [ctypes.c_uint32]
        L.po_log.restype = ctypes.c_uint16
        L.po_log.argtypes = [ctypes.c_uint16]
        L.po_input_log.restype = ctypes.c_uint16
        L.po_input_log.argtypes = [ctypes.c_uint16]
        L.po_coeff_log.restype = ctypes.c_uint16
        L.po_coeff_log.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.po_coeff.restype = ctypes.c_uint16
        L.po_coeff.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        L.po_cksum.restype = None
        L.po_cksum.argtypes = [_c_u8p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_uint, _c_u8p]
        L.po_cksum_verify.restype = ctypes.c_int
        L.po_cksum_verify.argtypes = [_c_u8p, ctypes.c_size_t, _c_u8p, ctypes.c_uint]
        L.po_muladd_region.restype = None
        L.po_muladd_region.argtypes = [_c_u8p, _c_u8p, ctypes.c_size_t, ctypes.c_uint16]
        L.po_recover.restype = ctypes.c_int
        L.po_recover.argtypes = [ctypes.c_uint, _c_u8p, ctypes.c_size_t, _c_szp, _c_u16p, _c_u16p,
                                 ctypes.c_uint, _c_u16p, ctypes.c_size_t, _c_u8p, ctypes.c_int]
        L.po_prng_fill_c.restype = None
        L.po_prng_fill_c.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                     _c_u8p, ctypes.c_size_t]
        L.po_prng_cksums.restype = None
        L.po_prng_cksums.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_size_t, ctypes.c_uint, _c_u8p,
                                     ctypes.c_int]
        L.po_md5.restype = None
        L.po_md5.argtypes = [_c_u8p, ctypes.c_size_t, ctypes.c_size_t, _c_u8p]
        L.po_crc32.restype = ctypes.c_uint32
        L.po_crc32.argtypes = [_c_u8p, ctypes.c_size_t, ctypes.c_size_t]
        L.po_init()
        _lib = L
    return _lib


def ref_available(portable=False):
    return os.path.exists(REF_PORTABLE_SO if portable else REF_SO)


_ref_portable = None


def ref(portable=False):
    """Load the reference harness (only present where /root/reference was built).
    portable=True loads the build that runs on any x86-64 host."""
    global _ref, _ref_portable
    if (_ref_portable if portable else _ref) is None:
        path = REF_PORTABLE_SO if portable else REF_SO
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (run `make -C oracle ref portable`)" % path)
        R = ctypes.CDLL(path)
        R.ref_info.restype = ctypes.c_int
        R.ref_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint),
                               ctypes.POINTER(ctypes.c_uint)]
        R.ref_gfmat_input_log.restype = ctypes.c_uint16
        R.ref_gfmat_input_log.argtypes = [ctypes.c_uint16]
        R.ref_gfmat_coeff.restype = ctypes.c_uint16
        R.ref_gfmat_coeff.argtypes = [ctypes.c_uint16, ctypes.c_uint16]
        R.ref_gf16_exp.restype = ctypes.c_uint16
        R.ref_gf16_exp.argtypes = [ctypes.c_uint16]
        R.ref_cksum_copy.restype = ctypes.c_int
        R.ref_cksum_copy.argtypes = [ctypes.c_int, _c_u8p, _c_u8p, ctypes.c_size_t, ctypes.c_size_t]
        R.ref_cksum_check.restype = ctypes.c_int
        R.ref_cksum_check.argtypes = [ctypes.c_int, _c_u8p, _c_u8p, ctypes.c_size_t]
        R.ref_method_cksum_size.restype = ctypes.c_int
        R.ref_method_cksum_size.argtypes = [ctypes.c_int]
        R.ref_recover.restype = ctypes.c_int
        R.ref_recover.argtypes = [ctypes.c_int, ctypes.c_uint, _c_u8p, ctypes.c_size_t, _c_szp, _c_u16p, _c_u16p,
                                  ctypes.c_uint, _c_u16p, ctypes.c_size_t, ctypes.c_size_t, _c_u8p, _c_intp,
                                  ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        if portable:
            _ref_portable = R
        else:
            _ref = R
    return _ref_portable if portable else _ref


# ---------------------------------------------------------------- helpers

def prng_bytes(seed, job, slice_idx, length, offset=0):
    """Bytes [offset, offset+length) of the synthetic stream of (seed, job, slice)."""
    out = np.empty(length, dtype=np.uint8)
    lib().po_prng_fill_c(seed, job, slice_idx, offset, _ptr(out), length)
    return out


def recover(inputs, in_lens, slice_size, exps=None, input_nums=None, coeffs=None, threads=0):
    """Oracle recovery: inputs is (nIn, stride) uint8, returns (nOut, slice_size) uint8.

    coeffs, if given, is (nIn, nOut) uint16 (explicit-coefficient path)."""
    L = lib()
    inputs = np.ascontiguousarray(inputs, dtype=np.uint8)
    n_in = inputs.shape[0]
    lens = np.ascontiguousarray(in_lens, dtype=np.uintp)
    if coeffs is not None:
        coeffs = np.ascontiguousarray(coeffs, dtype=np.uint16)
        n_out = coeffs.shape[1]
        ex = np.zeros(n_out, dtype=np.uint16)
        nums = np.zeros(max(n_in, 1), dtype=np.uint16)
    else:
        ex = np.ascontiguousarray(exps, dtype=np.uint16)
        n_out = ex.shape[0]
        nums = np.ascontiguousarray(input_nums if input_nums is not None else np.arange(n_in), dtype=np.uint16)
    out = np.zeros((n_out, slice_size), dtype=np.uint8)
    L.po_recover(n_in, _ptr(inputs), inputs.shape[1] if n_in else 0, _ptr(lens, _c_szp), _ptr(nums, _c_u16p),
                 _ptr(coeffs, _c_u16p), n_out, _ptr(ex, _c_u16p), slice_size, _ptr(out), threads)
    return out


def cksum(data, slice_len, width, src_len=None):
    """The gf16_cksum block copy_cksum stores after a slice of slice_len bytes."""
    L = lib()
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = data.shape[0] if src_len is None else src_len
    out = np.zeros(width, dtype=np.uint8)
    L.po_cksum(_ptr(data), n, slice_len, width, _ptr(out))
    return out


def prng_cksums(seed, job, count, slice_len, lens=None, width=64, first=0, threads=0):
    """gf16_cksum (count x width bytes) of synthetic slices first..first+count-1
    of (seed, job), lens[k] bytes each zero-padded to slice_len."""
    out = np.zeros((count, width), np.uint8)
    lp = None
    if lens is not None:
        la = np.ascontiguousarray(lens, dtype=np.uint32)
        assert la.size == count
        lp = la.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    lib().po_prng_cksums(seed, job, first, count, lp, slice_len, width, _ptr(out), threads)
    return out


def cksum_verify(data, ck, width):
    L = lib()
    data = np.ascontiguousarray(data, dtype=np.uint8)
    ck = np.ascontiguousarray(ck, dtype=np.uint8)
    return bool(L.po_cksum_verify(_ptr(data), data.shape[0], _ptr(ck), width))


def ref_recover(inputs, in_lens, slice_size, exps, input_nums=None, coeffs=None, split=0, threads=0, method=0,
                portable=False):
    """Run the reference PAR2Proc + PAR2ProcCPU; returns (outputs, valid, seconds)."""
    R = ref(portable)
    inputs = np.ascontiguousarray(inputs, dtype=np.uint8)
    n_in = inputs.shape[0]
    lens = np.ascontiguousarray(in_lens, dtype=np.uintp)
    ex = np.ascontiguousarray(exps, dtype=np.uint16)
    n_out = ex.shape[0]
    nums = np.ascontiguousarray(input_nums if input_nums is not None else np.arange(n_in), dtype=np.uint16)
    cf = None
    if coeffs is not None:
        cf = np.ascontiguousarray(coeffs, dtype=np.uint16)
    out = np.zeros((n_out, slice_size), dtype=np.uint8)
    valid = np.zeros(n_out, dtype=np.intc)
    secs = ctypes.c_double(0)
    rc = R.ref_recover(method, n_in, _ptr(inputs), inputs.shape[1] if n_in else 0, _ptr(lens, _c_szp),
                       _ptr(nums, _c_u16p), _ptr(cf, _c_u16p), n_out, _ptr(ex, _c_u16p), slice_size, split,
                       _ptr(out), _ptr(valid, _c_intp), threads, ctypes.byref(secs))
    if rc != 0:
        raise RuntimeError("ref_recover failed: %d" % rc)
    return out, valid.astype(bool), secs.value


def ref_info(portable=False):
    R = ref(portable)
    name = ctypes.create_string_buffer(64)
    ck = ctypes.c_uint(0)
    th = ctypes.c_uint(0)
    mid = R.ref_info(name, 64, ctypes.byref(ck), ctypes.byref(th))
    return {"method": name.value.decode(), "method_id": mid, "cksum_size": ck.value, "threads": th.value}


def ref_cksum_copy(data, slice_len, method=0, src_len=None):
    """Reference copy_cksum: returns (padded data, trailer) for a method (0 = AUTO)."""
    R = ref()
    data = np.ascontiguousarray(data, dtype=np.uint8)
    n = data.shape[0] if src_len is None else src_len
    width = R.ref_method_cksum_size(method)
    dst = np.zeros(slice_len + width + 64, dtype=np.uint8)
    w = R.ref_cksum_copy(method, _ptr(dst), _ptr(data), n, slice_len)
    return dst[:slice_len].copy(), dst[slice_len:slice_len + w].copy()
