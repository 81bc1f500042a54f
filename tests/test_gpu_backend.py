"""Backend-layer parity: PAR2ProcHIP (IPAR2ProcBackend over the C ABI) on every
golden case of the reference's PAR2ProcCPU, driven like ParPar's controller."""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P
from golden_io import case_inputs, manifest, small_outputs

pytestmark = pytest.mark.gpu
G = manifest()
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)


FORMS = {"plain": parpar_amd.FORM_PLAIN, "bilinear": parpar_amd.FORM_BILINEAR, "auto": parpar_amd.FORM_AUTO}


def new_backend(slice_size, batch, areas=2, ck=64, form="plain"):
    be = P.PAR2ProcHIP(0, areas)
    be.setSliceSize(slice_size)
    be.init(cksum_size=ck, input_batch=batch, form=FORMS[form])
    return be


def run_case(case, batch, areas, form="plain"):
    inp = case_inputs(case)
    S = case["S"]
    c = P.PAR2Proc()
    if case["split"]:
        parts = [(0, case["split"]), (case["split"], S - case["split"])]
    else:
        parts = [(0, S)]
    bes = []
    allocs = []
    for off, size in parts:
        be = P.PAR2ProcHIP(0, areas)
        bes.append(be)
        allocs.append((be, off, size))
    assert c.init(S, allocs)
    for be, off, size in allocs:
        be.init(cksum_size=64, input_batch=batch, form=FORMS[form])
    assert c.setRecoverySlices(np.array(case["exps"], np.uint16))
    coeffs = np.array(case["coeffs"], np.uint16) if case["coeffs"] is not None else None
    for k in range(case["n"]):
        if c.canAdd() in (parpar_amd.PPH_ADD_FULL, 3):
            c.waitForAdd()
        if coeffs is not None:
            c.addInput(inp[k], case["lens"][k], coeffs=coeffs[k])
        else:
            c.addInput(inp[k], case["lens"][k], inputNum=case["nums"][k])
    c.endInput()
    outs = np.zeros((len(case["exps"]) if coeffs is None else coeffs.shape[1], S), np.uint8)
    valid = []
    for r in range(outs.shape[0]):
        o = np.zeros(S, np.uint8)
        valid.append(c.getOutput(r, o))
        outs[r] = o
    for be in bes:
        be.close()
    return outs, valid


@pytest.mark.parametrize("batch,areas,form", [(0, 2, "plain"), (3, 2, "plain"), (1, 3, "plain"), (5, 1, "plain"),
                                              (0, 2, "bilinear"), (3, 2, "bilinear")])
def test_golden_small_cases(batch, areas, form):
    say("small cases batch", batch, "areas", areas, form)
    outs = small_outputs()
    for case in G["small_cases"]:
        say(" case", case["id"], "n", case["n"], "S", case["S"], "split", case["split"])
        got, valid = run_case(case, batch, areas, form)
        exp = outs["out_%d" % case["id"]]
        assert (got == exp).all(), case["id"]
        # checksum flags equal PAR2ProcCPU's, including its odd sub-slice case
        # where the product's high byte spills past the slice end
        assert list(valid) == list(case["valid"]), case["id"]


@pytest.mark.parametrize("form", ["plain", "bilinear"])
def test_config0_backend_golden(form):
    say("config0 through PAR2ProcHIP", form)
    c = G["config0"]
    S, n = c["slice"], c["n"]
    be = new_backend(S, 16 if form == "plain" else 64, form=form)
    be.setRecoverySlices(np.array(c["exps"], np.uint16))
    for i in range(n):
        if be.canAdd() == parpar_amd.PPH_ADD_FULL:
            be.waitForAdd()
        be.addInput(O.prng_bytes(0, 0, i, S), S, inputNum=i)
    be.endInput()
    for r in range(len(c["exps"])):
        o = np.zeros(S, np.uint8)
        assert be.getOutput(r, o)
        assert hashlib.sha256(o.tobytes()).hexdigest() == c["sha256"][r]
    be.close()


def test_empty_and_discard_and_corruption():
    say("empty / discard / corrupt")
    S = 4096
    be = new_backend(S, 4)
    be.setRecoverySlices(np.array([0, 1, 2], np.uint16))
    o = np.full(S, 7, np.uint8)
    # nothing added: zero fill, valid (controller.cpp:376-382)
    assert be.getOutput(1, o) and not o.any()
    inp = [O.prng_bytes(21, 0, k, S) for k in range(6)]
    for k in range(6):
        be.addInput(inp[k], S, inputNum=k)
    be.endInput()
    ref = O.recover(np.stack(inp), [S] * 6, S, [0, 1, 2])
    for r in range(3):
        assert be.getOutput(r, o) and (o == ref[r]).all()
    # discardOutput: the next batch overwrites instead of accumulating
    be.discardOutput()
    for k in range(2):
        be.addInput(inp[k], S, inputNum=k)
    be.endInput()
    ref2 = O.recover(np.stack(inp[:2]), [S] * 2, S, [0, 1, 2])
    for r in range(3):
        assert be.getOutput(r, o) and (o == ref2[r]).all()
    # host-forced corruption of a recovery slice is caught by gf16_cksum
    be.debug_corrupt_output(1, 100, 0x10)
    assert be.getOutput(0, o)
    assert not be.getOutput(1, o)
    be.close()


def test_current_slice_size_and_default_exponent():
    say("setCurrentSliceSize / default exponents")
    S = 8192
    be = new_backend(S, 4)
    be.setRecoverySlices(None, n=2)  # exponent 1 for every output (controller_cpu.cpp:176)
    inp = [O.prng_bytes(22, 0, k, S) for k in range(5)]
    for k in range(5):
        be.addInput(inp[k], S, inputNum=k)
    be.endInput()
    ref = O.recover(np.stack(inp), [S] * 5, S, [1, 1])
    o = np.zeros(S, np.uint8)
    assert be.getOutput(1, o) and (o == ref[1]).all()
    # shrink the slice for a second pass over the tail of the data
    be.discardOutput()
    T = 2050
    assert be.setCurrentSliceSize(T)
    be.setRecoverySlices(np.array([3, 9], np.uint16))
    for k in range(5):
        be.addInput(inp[k][:T], T, inputNum=k)
    be.endInput()
    ref = O.recover(np.stack([x[:T] for x in inp]), [T] * 5, T, [3, 9])
    o = np.zeros(T, np.uint8)
    for r in range(2):
        assert be.getOutput(r, o) and (o == ref[r]).all()
    # grow beyond the allocation
    U = 3 * S
    assert be.setCurrentSliceSize(U)
    be.setRecoverySlices(np.array([5], np.uint16))
    big = [O.prng_bytes(23, 0, k, U) for k in range(3)]
    for k in range(3):
        be.addInput(big[k], U, inputNum=k)
    be.endInput()
    o = np.zeros(U, np.uint8)
    assert be.getOutput(0, o) and (o == O.recover(np.stack(big), [U] * 3, U, [5])[0]).all()
    be.close()


@pytest.mark.parametrize("form", ["plain", "bilinear"])
def test_fill_and_dummy_inputs(form):
    say("fillInput / dummyInput", form)
    S = 4096
    be = new_backend(S, 2, areas=2, form=form)
    be.setRecoverySlices(np.array([1, 7], np.uint16))
    data = [O.prng_bytes(24, 0, k, S) for k in range(4)]
    done = False
    k = 0
    while not done:
        done = be.fillInput(data[k])
        k += 1
    assert k == 4
    for i in range(4):
        be.dummyInput(i)
    be.endInput()
    ref = O.recover(np.stack(data), [S] * 4, S, [1, 7])
    o = np.zeros(S, np.uint8)
    for r in range(2):
        assert be.getOutput(r, o) and (o == ref[r]).all()
    be.close()


def test_auto_form_large_batches():
    """PPH_FORM_AUTO: batches of 512 inputs with 64 recovery slices take the
    bilinear form, 8 recovery slices the plain one; both give the oracle's
    bytes.  Slots that fillInput prepared for one form refuse a recovery count
    that would switch to the other."""
    S, n = 8192, 1100
    data = np.stack([O.prng_bytes(25, 0, k, S) for k in range(n)])
    for R in (64, 8):
        say("auto form", R)
        be = new_backend(S, 512, areas=2, form="auto")
        exps = np.arange(3, 3 + R, dtype=np.uint16)
        be.setRecoverySlices(exps)
        for k in range(n):
            be.addInput(data[k], S if k < n - 1 else 3000, inputNum=k)
        be.endInput()
        lens = [S] * (n - 1) + [3000]
        padded = data.copy()
        padded[-1, 3000:] = 0
        ref = O.recover(padded, lens, S, exps)
        o = np.zeros(S, np.uint8)
        for r in range(R):
            assert be.getOutput(r, o) and (o == ref[r]).all(), (R, r)
        be.close()
    be = new_backend(S, 512, areas=1, form="auto")
    be.setRecoverySlices(np.arange(64, dtype=np.uint16))
    be.fillInput(data[0])
    with pytest.raises(parpar_amd.PPHError):
        be.setRecoverySlices(np.arange(8, dtype=np.uint16))
    be.close()


@pytest.mark.parametrize("ck", [8, 16, 32])
def test_other_checksum_widths(ck):
    say("cksum width", ck)
    S = 10002
    be = new_backend(S, 3, ck=ck)
    be.setRecoverySlices(np.array([0, 5, 65534], np.uint16))
    inp = [O.prng_bytes(25, ck, k, S) for k in range(7)]
    lens = [S] * 7
    lens[3] = 777
    for k in range(7):
        be.addInput(inp[k], lens[k], inputNum=k)
    be.endInput()
    ref_in = np.stack(inp)
    ref_in[3, 777:] = 0
    ref = O.recover(ref_in, lens, S, [0, 5, 65534])
    o = np.zeros(S, np.uint8)
    for r in range(3):
        assert be.getOutput(r, o) and (o == ref[r]).all()
    be.close()


def test_cpp_backend_driver():
    """The C++ PAR2ProcHIP class (controller_hip.cpp) driven through its futures."""
    exe = os.path.join(ROOT, "tests", "cpp", "par2proc_hip_test")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/par2proc_hip_test is not built (run __graft_entry__.build())")
    say("cpp driver")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "ALL OK" in r.stdout


def test_reference_controller_drives_par2prochip():
    """ParPar's own PAR2Proc (compiled from the reference in the build container)
    with PAR2ProcHIP as its only backend, end to end on the GPU."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_integration")
    if not os.path.exists(exe):
        pytest.skip("reference controller build not shipped with this tree")
    say("reference PAR2Proc + PAR2ProcHIP")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "ok=1" in r.stdout


def test_host_copy_corruption_is_reported():
    """getOutput's flag is computed on the host over the bytes it delivers: a
    byte flipped in host memory between the device-to-host copy and the copy to
    the caller makes that slice invalid, while the device copy stays intact."""
    say("host-side corruption of the copy out")
    S = 70000
    be = new_backend(S, 4)
    exps = np.array([0, 3, 11], np.uint16)
    be.setRecoverySlices(exps)
    inp = [O.prng_bytes(26, 0, k, S) for k in range(9)]
    for k in range(9):
        be.addInput(inp[k], S, inputNum=k)
    be.endInput()
    ref = O.recover(np.stack(inp), [S] * 9, S, exps)
    o = np.zeros(S, np.uint8)
    assert be.getOutput(0, o) and (o == ref[0]).all()
    be.debug_corrupt_host_copy(12345, 0x40)
    assert not be.getOutput(1, o)
    assert o[12345] == ref[1][12345] ^ 0x40
    # the next copies are clean again, out of order too (no prefetch hit)
    assert be.getOutput(2, o) and (o == ref[2]).all()
    assert be.getOutput(1, o) and (o == ref[1]).all()
    # the flag covers the last, partial checksum block
    be.debug_corrupt_host_copy(S - 1, 0x01)
    assert not be.getOutput(0, o)
    be.close()


@pytest.mark.parametrize("form", ["plain", "bilinear"])
def test_failed_batch_reports_every_output_invalid(form):
    """A batch whose submission fails loses its inputs; the backend keeps the
    error, and every getOutput of that computation returns False (not a valid
    flag over incomplete recovery data), until discardOutput starts afresh."""
    say("failed batch", form)
    S = 8192
    be = new_backend(S, 3, form=form)
    exps = np.array([1, 2, 9], np.uint16)
    be.setRecoverySlices(exps)
    inp = [O.prng_bytes(27, 0, k, S) for k in range(8)]
    o = np.zeros(S, np.uint8)
    for rnd in range(2):
        if rnd == 0:
            be.debug_fail_next_batch(parpar_amd.PPH_E_NOMEM)
        for k in range(8):
            if be.canAdd() == parpar_amd.PPH_ADD_FULL:
                be.waitForAdd()
            try:
                be.addInput(inp[k], S, inputNum=k)
            except parpar_amd.PPHError as e:
                assert rnd == 0 and e.status == parpar_amd.PPH_E_NOMEM
        be.endInput()
        if rnd == 0:
            assert be.last_status == parpar_amd.PPH_E_NOMEM
            for r in range(3):
                assert not be.getOutput(r, o)
        else:
            ref = O.recover(np.stack(inp), [S] * 8, S, exps)
            for r in range(3):
                assert be.getOutput(r, o) and (o == ref[r]).all()
        be.discardOutput()
    be.close()


def test_bilinear_workspace_shortage_runs_plain():
    """When the bilinear workspace cannot be allocated, a batch runs in the
    plain kernel instead: the same bytes, every flag valid."""
    say("bilinear workspace shortage")
    lib = parpar_amd.lib()
    S, n = 8192, 600
    data = np.stack([O.prng_bytes(28, 0, k, S) for k in range(n)])
    exps = np.arange(5, 5 + 70, dtype=np.uint16)
    lib.pph_bilinear_release()
    assert lib.pph_debug_bilinear_workspace_limit(1 << 20) == 0
    try:
        be = new_backend(S, 256, form="bilinear")
        be.setRecoverySlices(exps)
        for k in range(n):
            be.addInput(data[k], S, inputNum=k)
        be.endInput()
        ref = O.recover(data, [S] * n, S, exps)
        o = np.zeros(S, np.uint8)
        for r in range(len(exps)):
            assert be.getOutput(r, o) and (o == ref[r]).all(), r
        be.close()
    finally:
        lib.pph_debug_bilinear_workspace_limit(0)


def test_async_abi_explicit_areas():
    """The asynchronous entry points (explicit area and slot, submit with a
    completion function, getOutput on the worker thread) give the oracle's
    bytes; completions arrive once per batch and per output."""
    import ctypes
    import threading
    say("async ABI")
    lib = parpar_amd.lib()
    S, n, batch = 20000, 11, 4
    exps = np.array([0, 7, 300], np.uint16)
    be = new_backend(S, batch, areas=2)
    be.setRecoverySlices(exps)
    inp = [O.prng_bytes(29, 0, k, S) for k in range(n)]
    lens = [S] * n
    lens[-1] = 999
    inp[-1][999:] = 0
    assert lib.pph_proc_staging_areas(be._p) == 2
    done = []
    cv = threading.Condition()
    DONE = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int)
    OUT = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_int)

    @DONE
    def on_batch(user, status):
        with cv:
            done.append(("batch", user or 0, status))
            cv.notify_all()

    @OUT
    def on_output(user, status, valid):
        with cv:
            done.append(("out", user or 0, status, valid))
            cv.notify_all()

    area, slot, batches, first = 0, 0, 0, 1
    for k in range(n):
        buf = np.ascontiguousarray(inp[k])
        assert lib.pph_proc_stage_input(be._p, area, slot, buf.ctypes.data, lens[k], k, None) == 0
        slot += 1
        if slot == batch or k == n - 1:
            assert lib.pph_proc_submit(be._p, area, slot, first, ctypes.cast(on_batch, ctypes.c_void_p), batches) == 0
            first, slot, batches = 0, 0, batches + 1
            area ^= 1
            with cv:  # the next area is free once its last batch is done
                cv.wait_for(lambda: sum(1 for d in done if d[0] == "batch") >= batches - 1, timeout=60)
    with cv:
        assert cv.wait_for(lambda: sum(1 for d in done if d[0] == "batch") == batches, timeout=60)
    assert all(d[2] == 0 for d in done)
    outs = [np.zeros(S, np.uint8) for _ in exps]
    for r in range(len(exps)):
        assert lib.pph_proc_get_output_async(be._p, r, outs[r].ctypes.data, ctypes.cast(on_output, ctypes.c_void_p),
                                             100 + r) == 0
    assert lib.pph_proc_drain(be._p) == 0
    got = [d for d in done if d[0] == "out"]
    assert [d[1] for d in got] == [100, 101, 102]
    assert all(d[2] == 0 and d[3] == 1 for d in got)
    ref = O.recover(np.stack(inp), lens, S, exps)
    for r in range(len(exps)):
        assert (outs[r] == ref[r]).all(), r
    be.close()


@pytest.mark.parametrize("mode,args", [("hip", ["65536", "300", "20", "64"]),
                                       ("hip", ["4098", "37", "5", "4"]),
                                       ("split", ["65536", "120", "12", "16", "20000"]),
                                       ("split", ["10000", "29", "3", "5", "4002"])])
def test_reference_libuv_controller_drives_par2prochip(mode, args):
    """ParPar's PAR2Proc in its libuv build (the form of ParPar's Node binding)
    on a uv loop with PAR2ProcHIP's libuv form as a backend, alone or splitting
    each slice with ParPar's own PAR2ProcCPU: every recovery slice byte-exact
    against the oracle, every flag valid, adds refused while busy retried from
    the progress callback."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_integration_uv")
    if not os.path.exists(exe):
        pytest.skip("reference controller build not shipped with this tree")
    say("reference libuv PAR2Proc", mode, args)
    r = subprocess.run([exe, mode] + args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "ALL OK" in r.stdout
