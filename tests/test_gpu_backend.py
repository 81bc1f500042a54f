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
