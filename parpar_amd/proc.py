"""Python mirror of the PAR2ProcHIP backend and device-buffer helpers.

``PAR2ProcHIP`` exposes the backend operations of ParPar's plugin interface
(IPAR2ProcBackend, gf16/controller.h:77-217) with the same names and argument
meaning, calling the C ABI of libparpar_hip.so.  ``PAR2Proc`` is a small
restatement of ParPar's controller (gf16/controller.cpp) for one or more HIP
backends splitting each slice by byte range, as the reference controller
splits between CPU and OpenCL backends (controller.cpp:13-84,243-264,335-399).
"""
import ctypes

import numpy as np

from . import PPH_ADD_FULL, PPH_ADD_OK, PPH_ADD_OK_BUSY, PPH_E_INVALID, PPH_E_STATE, FORM_AUTO, check, lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceBuffer:
    """A device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().pph_malloc(ctypes.byref(p), self.nbytes), "pph_malloc")
        self.ptr = p

    @property
    def addr(self):
        return self.ptr.value

    def offset(self, off):
        return ctypes.c_void_p(self.ptr.value + off)

    def upload(self, arr, off=0):
        arr = np.ascontiguousarray(arr)
        assert off + arr.nbytes <= self.nbytes
        check(lib().pph_memcpy_h2d(self.offset(off), _ptr(arr), arr.nbytes), "pph_memcpy_h2d")

    def download(self, nbytes=None, off=0, dtype=np.uint8):
        n = self.nbytes - off if nbytes is None else nbytes
        out = np.empty(n, dtype=np.uint8)
        check(lib().pph_memcpy_d2h(_ptr(out), self.offset(off), n), "pph_memcpy_d2h")
        return out.view(dtype)

    def zero(self):
        check(lib().pph_memset(self.ptr, 0, self.nbytes), "pph_memset")

    def close(self):
        if self.ptr is not None and self.ptr.value:
            lib().pph_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def slot_layout(slice_size, cksum_size=64):
    ck = ctypes.c_size_t()
    slot = ctypes.c_size_t()
    check(lib().pph_slot_layout(slice_size, cksum_size, ctypes.byref(ck), ctypes.byref(slot)), "pph_slot_layout")
    return ck.value, slot.value


class PAR2ProcHIP:
    """One GPU recovery backend (PAR2ProcOCL's role, gf16/controller_ocl.cpp)."""

    def __init__(self, device=0, staging_areas=2):
        p = ctypes.c_void_p()
        check(lib().pph_proc_create(device, staging_areas, ctypes.byref(p)), "pph_proc_create")
        self._p = p
        self._recovery = 0

    def close(self):
        if self._p is not None:
            lib().pph_proc_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- IPAR2ProcBackend
    def setSliceSize(self, size):
        check(lib().pph_proc_set_slice_size(self._p, size), "setSliceSize")

    def init(self, cksum_size=64, input_batch=0, form=FORM_AUTO):
        check(lib().pph_proc_init(self._p, cksum_size, input_batch, form), "init")
        return True

    def setCurrentSliceSize(self, size):
        check(lib().pph_proc_set_current_slice_size(self._p, size), "setCurrentSliceSize")
        return True

    def setRecoverySlices(self, exponents, n=None):
        if exponents is None:
            arr = None
            n = int(n)
        else:
            arr = np.ascontiguousarray(exponents, dtype=np.uint16)
            n = arr.shape[0]
        check(lib().pph_proc_set_recovery_slices(self._p, n, _ptr(arr)), "setRecoverySlices")
        self._recovery = n
        return True

    def getNumRecoverySlices(self):
        return lib().pph_proc_num_recovery_slices(self._p)

    def canAdd(self):
        return lib().pph_proc_can_add(self._p)

    def waitForAdd(self):
        check(lib().pph_proc_wait_for_add(self._p), "waitForAdd")

    def addInput(self, buffer, size=None, inputNum=0, flush=False, coeffs=None):
        buf = np.ascontiguousarray(buffer, dtype=np.uint8)
        n = buf.nbytes if size is None else int(size)
        if coeffs is not None:
            cf = np.ascontiguousarray(coeffs, dtype=np.uint16)
            check(lib().pph_proc_add_input_coeffs(self._p, _ptr(buf), n, _ptr(cf), int(flush)), "addInput(coeffs)")
        else:
            check(lib().pph_proc_add_input(self._p, _ptr(buf), n, int(inputNum), int(flush)), "addInput")

    def dummyInput(self, inputNum, flush=False):
        check(lib().pph_proc_dummy_input(self._p, inputNum, int(flush)), "dummyInput")

    def fillInput(self, buffer):
        buf = np.ascontiguousarray(buffer, dtype=np.uint8)
        return check(lib().pph_proc_fill_input(self._p, _ptr(buf)), "fillInput") == 1

    def flush(self):
        check(lib().pph_proc_flush(self._p), "flush")

    def endInput(self):
        """Waits for every submitted batch.  A failed batch does not raise here:
        its status is kept (last_status) and getOutput reports False."""
        st = lib().pph_proc_end_input(self._p)
        self.last_status = st
        if st in (PPH_E_INVALID, PPH_E_STATE):
            check(st, "endInput")

    def getOutput(self, index, output):
        """The checksum result of the slice copied into output; False also when
        a batch of this recovery computation failed (its status is kept in
        last_status), as the C++ PAR2ProcHIP reports it."""
        valid = ctypes.c_int(0)
        st = lib().pph_proc_get_output(self._p, index, _ptr(output), ctypes.byref(valid))
        self.last_status = st
        if st in (PPH_E_INVALID, PPH_E_STATE):
            check(st, "getOutput")
        return st == 0 and bool(valid.value)

    def discardOutput(self):
        lib().pph_proc_discard_output(self._p)

    def _hasAdded(self):
        return bool(lib().pph_proc_has_added(self._p))

    def isEmpty(self):
        return bool(lib().pph_proc_is_empty(self._p))

    def freeProcessingMem(self):
        check(lib().pph_proc_free_processing_mem(self._p), "freeProcessingMem")

    def getInputBatchSize(self):
        return lib().pph_proc_input_batch_size(self._p)

    def getBatchesStarted(self):
        return lib().pph_proc_batches_started(self._p)

    def setMinInputBatchSize(self, n):
        lib().pph_proc_set_min_input_batch_size(self._p, n)

    def debug_corrupt_output(self, index, offset, mask=1):
        check(lib().pph_proc_debug_corrupt_output(self._p, index, offset, mask), "debug_corrupt_output")

    def debug_corrupt_host_copy(self, offset, mask=1):
        check(lib().pph_proc_debug_corrupt_host_copy(self._p, offset, mask), "debug_corrupt_host_copy")

    def debug_fail_next_batch(self, status):
        check(lib().pph_proc_debug_fail_next_batch(self._p, status), "debug_fail_next_batch")


class PacketMd5:
    """Recovery-packet MD5 on the device: HasherOutput's MD5Multi
    (src/gf.cc:1070-1219) fed as lib/par2.js:243-262,300-327 feeds it.  Region
    r hashes set_id || "PAR 2.0\\0RecvSlic" || u32le(exps[r]) || the recovery
    slice, appended in pieces of even length."""

    def __init__(self, regions):
        h = ctypes.c_void_p()
        check(lib().pph_packet_md5_create(int(regions), ctypes.byref(h)), "pph_packet_md5_create")
        self._h = h
        self.regions = int(regions)
        self._out = DeviceBuffer(16 * self.regions)

    def begin(self, set_id, exps):
        sid = np.frombuffer(bytes(set_id), np.uint8)
        assert sid.size == 16
        e = np.ascontiguousarray(exps, dtype=np.uint16)
        assert e.size == self.regions
        check(lib().pph_packet_md5_begin(self._h, _ptr(sid), _ptr(e), None), "pph_packet_md5_begin")

    def update(self, slots, slot_bytes, length):
        """Append bytes [0, length) of each slot (DeviceBuffer, or None: zeros)."""
        ptr = slots.ptr if slots is not None else None
        check(lib().pph_packet_md5_update(self._h, ptr, slot_bytes, length, None), "pph_packet_md5_update")

    def update_from(self, proc):
        """Append the backend's current recovery slices."""
        check(lib().pph_proc_packet_md5_update(proc._p, self._h), "pph_proc_packet_md5_update")

    def final(self):
        check(lib().pph_packet_md5_final(self._h, self._out.ptr, None), "pph_packet_md5_final")
        check(lib().pph_synchronize(), "pph_synchronize")
        return self._out.download(16 * self.regions).reshape(self.regions, 16)

    def close(self):
        if self._h is not None:
            lib().pph_packet_md5_destroy(self._h)
            self._h = None
        self._out.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PAR2Proc:
    """ParPar's controller over backends that each own a byte range of every slice."""

    def __init__(self):
        self.backends = []
        self.has_added = False
        self.slice_size = 0

    def init(self, slice_size, allocs):
        """allocs: list of (backend, offset, size), covering [0, slice_size) without overlap."""
        self.slice_size = slice_size
        self.backends = []
        pos = 0
        for be, off, size in sorted(allocs, key=lambda a: a[1]):
            if off != pos or (off & 1) or ((size & 1) and off + size != slice_size):
                return False
            pos += size
            be.setSliceSize(size)
            self.backends.append([be, off, size])
        return pos == slice_size

    def setRecoverySlices(self, exponents):
        return all(be.setRecoverySlices(exponents) for be, _, _ in self.backends)

    def canAdd(self):
        states = [be.canAdd() for be, _, _ in self.backends]
        if all(s == PPH_ADD_FULL for s in states):
            return 3  # PROC_ADD_ALL_FULL
        if all(s == PPH_ADD_OK for s in states):
            return PPH_ADD_OK
        if any(s == PPH_ADD_FULL for s in states):
            return PPH_ADD_FULL
        return PPH_ADD_OK_BUSY

    def waitForAdd(self):
        for be, _, _ in self.backends:
            be.waitForAdd()

    def addInput(self, buffer, size, inputNum=0, flush=False, coeffs=None):
        buf = np.ascontiguousarray(buffer, dtype=np.uint8)
        for be, off, bsize in self.backends:
            if off >= size:
                continue
            amount = min(size - off, bsize)
            if amount == 0:
                continue
            be.addInput(buf[off:off + amount], amount, inputNum, flush, coeffs=coeffs)
        self.has_added = True

    def endInput(self):
        for be, _, size in self.backends:
            if size:
                be.endInput()

    def getOutput(self, index, output):
        if not self.has_added:
            output[:self.slice_size] = 0
            return True
        ok = True
        for be, off, size in self.backends:
            if not size:
                continue
            if not be._hasAdded():
                output[off:off + size] = 0
                continue
            part = np.empty(size, dtype=np.uint8)
            ok = be.getOutput(index, part) and ok
            output[off:off + size] = part
        return ok
