"""The C ABI library: it loads, exports every symbol include/parpar_hip.h
declares, and its host-side helpers agree with the golden vectors.  No compute
call needs a GPU here."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P
from golden_io import manifest

G = manifest()


def test_library_loads_and_exports_every_declared_symbol():
    L = parpar_amd.lib()
    names = parpar_amd.declared_symbols()
    assert len(names) > 40
    for n in names:
        assert hasattr(L, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", parpar_amd.LIB_PATH], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    # the ctypes signature table covers the whole header
    assert set(names) <= set(parpar_amd._SIGS)


def test_library_is_built_for_gfx950():
    data = open(parpar_amd.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_host_coefficient_helpers_match_reference_vectors():
    L = parpar_amd.lib()
    for k, v in G["input_log"].items():
        assert L.pph_input_log(int(k)) == v
    for k, v in G["coeff"].items():
        i, e = map(int, k.split(","))
        assert L.pph_coeff(i, e) == v
    OL = O.lib()
    rng = np.random.default_rng(2)
    for a, b in rng.integers(0, 65536, (5000, 2)):
        assert L.pph_gf_mul(int(a), int(b)) == OL.po_mul_slow(int(a), int(b))
    for i in range(0, 32768, 37):
        lg = L.pph_input_log(i)
        for e in (0, 1, 2, 65534, 40000):
            assert L.pph_coeff_log(lg, e) == OL.po_coeff_log(lg, e)


def test_slot_layout():
    for S in (0, 1, 2, 100, 1001, 65536, 1048576, 16 * 1048576 + 2):
        for ck in (8, 16, 32, 64):
            off, slot = P.slot_layout(S, ck)
            assert off >= S and off % 16 == 0 and off - S < 16
            assert slot % 1024 == 0 and slot >= off + ck and slot - (off + ck) < 1024
    with pytest.raises(parpar_amd.PPHError):
        P.slot_layout(100, 24)


def test_status_strings():
    L = parpar_amd.lib()
    for s in (0, -1, -2, -3, -4, -5):
        assert L.pph_strerror(s)


def test_no_device_reports_unsupported():
    if parpar_amd.device_count() > 0:
        pytest.skip("a HIP device is present")
    p = ctypes.c_void_p()
    assert parpar_amd.lib().pph_proc_create(0, 2, ctypes.byref(p)) == parpar_amd.PPH_E_UNSUPPORTED
    with pytest.raises(RuntimeError):
        parpar_amd.require_device()


def test_serve_config_validation():
    """pph_serve_create checks its settings before touching a device."""
    from parpar_amd import serve as SV
    L = parpar_amd.lib()
    h = ctypes.c_void_p()
    base = dict(hbm_budget=1 << 20, max_compute=2, md5_quantum=0, staging_bytes=0, bilinear_workspace=0, tuning=None,
                md5_host_threads=0, upload_threads=0)
    for bad in (dict(upload_threads=17), dict(md5_host_threads=257), dict(max_compute=0), dict(max_compute=65),
                dict(md5_quantum=100), dict(hbm_budget=0)):
        cfg = SV.ServeConfig(**dict(base, **bad))
        assert L.pph_serve_create(0, ctypes.byref(cfg), ctypes.byref(h)) == parpar_amd.PPH_E_INVALID, bad
    assert L.pph_serve_create(0, None, ctypes.byref(h)) == parpar_amd.PPH_E_INVALID
    if parpar_amd.device_count() == 0:
        cfg = SV.ServeConfig(**dict(base, md5_host_threads=4, upload_threads=16))
        assert L.pph_serve_create(0, ctypes.byref(cfg), ctypes.byref(h)) == parpar_amd.PPH_E_UNSUPPORTED


def test_controller_allocation_rules():
    # PAR2Proc::checkBackendAllocation (gf16/controller.cpp:42-84): even offsets,
    # full cover, odd size only for the last range
    class Dummy:
        def setSliceSize(self, s):
            self.s = s

    c = P.PAR2Proc()
    assert c.init(100, [(Dummy(), 0, 50), (Dummy(), 50, 50)])
    assert not c.init(100, [(Dummy(), 0, 51), (Dummy(), 51, 49)])
    assert not c.init(100, [(Dummy(), 0, 50), (Dummy(), 52, 48)])
    assert c.init(101, [(Dummy(), 0, 50), (Dummy(), 50, 51)])


REF_INTEGRATION = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                               "ref_integration")


@pytest.mark.skipif(not os.path.exists(REF_INTEGRATION), reason="reference controller build not present")
def test_par2prochip_links_with_reference_controller():
    # ParPar's own PAR2Proc (built from the reference sources) with PAR2ProcHIP
    # as its backend: the class compiles against controller.h and links
    r = subprocess.run([REF_INTEGRATION, "--construct-only"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert "constructed" in r.stdout


def test_bilinear_formula_matches_gf_multiply():
    """The rank-81 Karatsuba formula behind the bilinear dense form (host copy of
    its expansion / recombination maps) multiplies like GF(2^16): all 65536
    coefficients against the basis words and random pairs."""
    import numpy as np
    import oracle as O
    L, R = parpar_amd.lib(), O.lib()
    for c in range(0, 65536, 7):
        for b in (0, 3, 8, 15):
            assert L.pph_bilinear_mul(c, 1 << b) == R.po_mul(c, 1 << b)
    rng = np.random.default_rng(81)
    for c, x in rng.integers(0, 65536, (20000, 2)):
        assert L.pph_bilinear_mul(int(c), int(x)) == R.po_mul(int(c), int(x))


def _uv_harness():
    import os
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "ref_integration_uv")
    if not os.path.exists(exe):
        pytest.skip("reference controller build not present (built where the reference sources are)")
    return exe


def test_libuv_form_builds_against_the_reference_controller():
    """PAR2ProcHIP's libuv form compiles against ParPar's own gf16/controller.h
    with USE_LIBUV and links with ParPar's PAR2Proc; without a device it
    constructs, reports no device and closes its loop handles."""
    import subprocess
    r = subprocess.run([_uv_harness(), "--construct-only"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert "constructed" in r.stdout


@pytest.mark.parametrize("args", [["65536", "120", "12", "16"], ["10000", "29", "3", "5"]])
def test_libuv_harness_with_the_reference_cpu_backend(args):
    """The uv-loop harness of the libuv integration test, run with ParPar's own
    PAR2ProcCPU as the only backend: it drives the controller the way
    lib/par2.js does (refused adds retried from the progress callback) and its
    results equal the oracle's."""
    import subprocess
    r = subprocess.run([_uv_harness(), "cpu"] + args, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout
    assert "ALL OK" in r.stdout
