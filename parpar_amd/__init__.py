"""parpar_amd — MI355X-native GF(2^16) recovery engine for ParPar.

The product is libparpar_hip.so (HIP kernels for gfx950 + C ABI, declared in
include/parpar_hip.h) and the C++ PAR2ProcHIP backend compiled into it.  This
package loads that library with ctypes and mirrors the backend interface for
Python callers (``parpar_amd.proc``).  There is no CPU fallback: without the
built library, or without a HIP device, the compute entry points raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libparpar_hip.so")

PPH_OK = 0
PPH_E_INVALID = -1
PPH_E_NOMEM = -2
PPH_E_STATE = -3
PPH_E_DEVICE = -4
PPH_E_UNSUPPORTED = -5

PPH_ADD_OK = 0
PPH_ADD_OK_BUSY = 1
PPH_ADD_FULL = 2

FORM_AUTO = 0
FORM_PLAIN = 1
FORM_DENSE = 2
FORM_BILINEAR = 3

_lib = None

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u = ctypes.c_uint
_i = ctypes.c_int
_u16 = ctypes.c_uint16
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64

# name -> (restype, argtypes)
_SIGS = {
    "pph_strerror": (ctypes.c_char_p, [_i]),
    "pph_last_hip_error": (_i, []),
    "pph_device_count": (_i, []),
    "pph_device_info": (_i, [_i, ctypes.c_char_p, _sz, ctypes.POINTER(_i)]),
    "pph_device_memory": (_i, [_i, ctypes.POINTER(_sz)]),
    "pph_input_log": (_u16, [_u16]),
    "pph_coeff_log": (_u16, [_u16, _u16]),
    "pph_coeff": (_u16, [_u16, _u16]),
    "pph_gf_mul": (_u16, [_u16, _u16]),
    "pph_slot_layout": (_i, [_sz, _u, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "pph_proc_create": (_i, [_i, _u, ctypes.POINTER(_vp)]),
    "pph_proc_init": (_i, [_vp, _u, _u, _i]),
    "pph_proc_destroy": (None, [_vp]),
    "pph_proc_set_slice_size": (_i, [_vp, _sz]),
    "pph_proc_set_current_slice_size": (_i, [_vp, _sz]),
    "pph_proc_set_recovery_slices": (_i, [_vp, _u, _vp]),
    "pph_proc_can_add": (_i, [_vp]),
    "pph_proc_wait_for_add": (_i, [_vp]),
    "pph_proc_add_input": (_i, [_vp, _vp, _sz, _u16, _i]),
    "pph_proc_add_input_coeffs": (_i, [_vp, _vp, _sz, _vp, _i]),
    "pph_proc_dummy_input": (_i, [_vp, _u16, _i]),
    "pph_proc_fill_input": (_i, [_vp, _vp]),
    "pph_proc_flush": (_i, [_vp]),
    "pph_proc_end_input": (_i, [_vp]),
    "pph_proc_get_output": (_i, [_vp, _u, _vp, ctypes.POINTER(_i)]),
    "pph_proc_discard_output": (None, [_vp]),
    "pph_proc_has_added": (_i, [_vp]),
    "pph_proc_is_empty": (_i, [_vp]),
    "pph_proc_free_processing_mem": (_i, [_vp]),
    "pph_proc_input_batch_size": (_u, [_vp]),
    "pph_proc_batches_started": (_u, [_vp]),
    "pph_proc_set_min_input_batch_size": (None, [_vp, _u]),
    "pph_proc_num_recovery_slices": (_i, [_vp]),
    "pph_proc_debug_corrupt_output": (_i, [_vp, _u, _sz, ctypes.c_uint8]),
    "pph_proc_debug_corrupt_host_copy": (_i, [_vp, _sz, ctypes.c_uint8]),
    "pph_proc_debug_fail_next_batch": (_i, [_vp, _i]),
    "pph_debug_bilinear_workspace_limit": (_i, [_sz]),
    "pph_muladd_bilinear_noalloc": (_i, [_vp, _sz, _u, _vp, _vp, _vp, _u, _i, _vp]),
    "pph_proc_staging_areas": (_u, [_vp]),
    "pph_proc_stage_input": (_i, [_vp, _u, _u, _vp, _sz, _u16, _vp]),
    "pph_proc_stage_dummy": (_i, [_vp, _u, _u, _u16]),
    "pph_proc_fill_slot": (_i, [_vp, _u, _u, _vp]),
    "pph_proc_submit": (_i, [_vp, _u, _u, _i, _vp, _vp]),
    "pph_proc_get_output_async": (_i, [_vp, _u, _vp, _vp, _vp]),
    "pph_proc_drain": (_i, [_vp]),
    "pph_set_device": (_i, [_i]),
    "pph_mem_get_info": (_i, [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "pph_malloc": (_i, [ctypes.POINTER(_vp), _sz]),
    "pph_free": (_i, [_vp]),
    "pph_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "pph_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "pph_memset": (_i, [_vp, _i, _sz]),
    "pph_copy": (_i, [_vp, _vp, _sz, _vp]),
    "pph_lds_gather_ceiling": (_i, [_u, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)]),
    "pph_synchronize": (_i, []),
    "pph_stream_create": (_i, [ctypes.POINTER(_vp)]),
    "pph_stream_destroy": (_i, [_vp]),
    "pph_stream_synchronize": (_i, [_vp]),
    "pph_timer_create": (_i, [ctypes.POINTER(_vp)]),
    "pph_timer_destroy": (_i, [_vp]),
    "pph_timer_start": (_i, [_vp, _vp]),
    "pph_timer_stop": (_i, [_vp, _vp]),
    "pph_timer_elapsed_ms": (_i, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "pph_fill_random": (_i, [_vp, _sz, _u, _sz, _u64, _u32, _u32, _vp]),
    "pph_prep_slots": (_i, [_vp, _sz, _u, _vp, _sz, _u, _i, _vp]),
    "pph_prep_slots_from": (_i, [_vp, _vp, _sz, _u, _vp, _sz, _u, _i, _vp]),
    "pph_muladd_plain": (_i, [_vp, _sz, _u, _vp, _vp, _vp, _u, _vp, _u, _i, _u, _vp]),
    "pph_muladd_dense": (_i, [_vp, _sz, _u, _vp, _vp, _vp, _u, _i, _vp]),
    "pph_autotune": (_i, [_vp]),
    "pph_batch_create": (_i, [_vp, _u, _u, _vp, ctypes.POINTER(_vp)]),
    "pph_batch_plan": (_i, [_vp, _u, ctypes.POINTER(_i), ctypes.POINTER(_u), ctypes.POINTER(ctypes.c_double)]),
    "pph_batch_run": (_i, [_vp, _i, _vp]),
    "pph_batch_destroy": (None, [_vp]),
    "pph_batch_fused_ms": (_i, [_vp, ctypes.POINTER(ctypes.c_float)]),
    "pph_slice_hash": (_i, [_vp, _sz, _u, _vp, _sz, _vp, _vp, _vp]),
    "pph_file_md5": (_i, [_vp, _sz, _u, _sz, _u64, _vp, _vp]),
    "pph_slice_hash_ifsc": (_i, [_vp, _sz, _u, _vp, _sz, _vp, _vp]),
    "pph_packet_md5_create": (_i, [_u, _vp]),
    "pph_packet_md5_destroy": (None, [_vp]),
    "pph_packet_md5_begin": (_i, [_vp, _vp, _vp, _vp]),
    "pph_packet_md5_update": (_i, [_vp, _vp, _sz, _sz, _vp]),
    "pph_packet_md5_final": (_i, [_vp, _vp, _vp]),
    "pph_proc_packet_md5_update": (_i, [_vp, _vp]),
    "pph_verify_slots": (_i, [_vp, _sz, _u, _sz, _u, _vp, _vp]),
    "pph_muladd_bilinear": (_i, [_vp, _sz, _u, _vp, _vp, _vp, _u, _i, _vp]),
    "pph_muladd_bilinear_coef": (_i, [_vp, _sz, _u, _vp, _u, _vp, _u, _i, _vp]),
    "pph_bilinear_workspace": (_i, [_sz, _u, _u, ctypes.POINTER(_sz)]),
    "pph_bilinear_release": (_i, []),
    "pph_bilinear_mul": (_u16, [_u16, _u16]),
    "pph_serve_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "pph_serve_submit": (_i, [_vp, _vp, ctypes.POINTER(_u64)]),
    "pph_serve_poll": (_i, [_vp, _vp, _u, ctypes.POINTER(_u), _i]),
    "pph_serve_drain": (_i, [_vp, _i]),
    "pph_serve_get_stats": (_i, [_vp, _vp]),
    "pph_serve_destroy": (None, [_vp]),
}


class PPHError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = "%s failed: %s (%d)" % (what, lib().pph_strerror(status).decode(), status)
        if status == PPH_E_DEVICE:
            msg += " [hip error %d]" % lib().pph_last_hip_error()
        super().__init__(msg)


def lib():
    """The loaded libparpar_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libparpar_hip.so is not built: run `python -m parpar_amd.build` "
                               "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, what=""):
    if status < 0:
        raise PPHError(status, what)
    return status


def declared_symbols():
    """Names of every function include/parpar_hip.h declares."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "parpar_hip.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(pph_[a-z0-9_]+)\s*\(", text)))


def device_count():
    return lib().pph_device_count()


def require_device():
    n = device_count()
    if n <= 0:
        raise RuntimeError("no HIP device visible: the parpar_amd engine has no CPU fallback")
    return n
