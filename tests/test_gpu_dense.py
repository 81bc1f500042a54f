"""Dense forms' parity: the int8-MFMA bit-matrix contraction and the bilinear
(Karatsuba, 81 GEMMs on fp4 MFMA) form against the oracle, the reference's golden
vectors and the plain kernel.  Calls go through the C ABI."""
import hashlib

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P
from golden_io import expected_valid, manifest

pytestmark = pytest.mark.gpu
G = manifest()


def say(*a):
    print(*a, flush=True)


def lib():
    return parpar_amd.lib()


FORMS = ["dense", "bilinear"]


def run_dense(datas, lens, S, exps, nums=None, ck=64, split=None, also_plain=False, form="dense"):
    n = len(datas)
    _, slot = P.slot_layout(S, ck)
    assert slot % 1024 == 0
    host = np.full((max(n, 1), slot), 0xC3, np.uint8)
    for k, d in enumerate(datas):
        host[k, :len(d)] = d
    src = P.DeviceBuffer(host.nbytes)
    src.upload(host)
    d_lens = P.DeviceBuffer(4 * max(n, 1))
    d_lens.upload(np.array(lens or [0], np.uint32))
    nums = list(range(n)) if nums is None else list(nums)
    d_logs = P.DeviceBuffer(2 * max(n, 1))
    d_logs.upload(np.array([lib().pph_input_log(x) for x in nums] or [0], np.uint16))
    n_out = len(exps)
    d_exps = P.DeviceBuffer(2 * n_out)
    d_exps.upload(np.array(exps, np.uint16))
    data = P.DeviceBuffer(host.nbytes)
    parpar_amd.check(lib().pph_prep_slots_from(data.ptr, src.ptr, slot, n, d_lens.ptr, S, ck, 0, None), "prep")
    out = P.DeviceBuffer(slot * n_out)
    out.upload(np.full(slot * n_out, 0x77, np.uint8))
    parts = [(0, n)] if split is None else [(0, split), (split, n)]
    for bi, (a, b) in enumerate(parts):
        fn = lib().pph_muladd_dense if form == "dense" else lib().pph_muladd_bilinear
        parpar_amd.check(fn(data.offset(a * slot), slot, b - a, d_logs.offset(2 * a), d_exps.ptr,
                            out.ptr, n_out, 1 if bi else 0, None), form)
    valid = P.DeviceBuffer(4 * n_out)
    parpar_amd.check(lib().pph_verify_slots(out.ptr, slot, n_out, S, ck, valid.ptr, None), "verify")
    parpar_amd.check(lib().pph_synchronize())
    res = out.download().reshape(n_out, slot)
    v = valid.download(4 * n_out).view(np.int32).copy()
    plain = None
    if also_plain:
        logd = P.DeviceBuffer(host.nbytes)
        parpar_amd.check(lib().pph_prep_slots_from(logd.ptr, src.ptr, slot, n, d_lens.ptr, S, ck, 1, None))
        pout = P.DeviceBuffer(slot * n_out)
        parpar_amd.check(lib().pph_muladd_plain(logd.ptr, slot, n, d_logs.ptr, d_exps.ptr, None, 0, pout.ptr, n_out,
                                                0, 0, None))
        parpar_amd.check(lib().pph_synchronize())
        plain = pout.download().reshape(n_out, slot)
    return res, v, plain


@pytest.mark.parametrize("form", FORMS)
def test_dense_single_input_single_slice(form):
    say(form, "1x1")
    S = 2048
    d = O.prng_bytes(70, 0, 0, S)
    got, valid, _ = run_dense([d], [S], S, [1], form=form)
    ref = O.recover(d[None, :], [S], S, [1])
    assert (got[0, :S] == ref[0]).all()
    assert valid.all()


@pytest.mark.parametrize("n,n_out,S", [(1, 1, 1024), (2, 8, 4096), (3, 9, 4000), (17, 13, 8192), (64, 40, 20000),
                                       (5, 3, 1), (300, 70, 6000)])
@pytest.mark.parametrize("form", FORMS)
def test_dense_vs_oracle(n, n_out, S, form):
    say(form, n, n_out, S)
    rng = np.random.default_rng(n * 1000 + n_out)
    lens = [S] * n
    lens[-1] = int(rng.integers(0, S + 1))
    datas = [O.prng_bytes(71, n, k, lens[k]) for k in range(n)]
    if lens[0] > 100:
        datas[0][20:60] = 0
    exps = rng.integers(0, 65535, n_out).astype(np.uint16)
    exps[0] = 0
    nums = rng.integers(0, 32768, n).astype(np.uint16)
    got, valid, plain = run_dense(datas, lens, S, exps, nums=nums, split=n // 2 if n > 2 else None,
                                  also_plain=True, form=form)
    inp = np.zeros((n, max(S, 1)), np.uint8)
    for k in range(n):
        inp[k, :lens[k]] = datas[k]
    ref = O.recover(inp, lens, S, exps, input_nums=nums)
    assert (got[:, :S] == ref).all()
    assert ((valid == 1) == expected_valid(inp, lens, S, exps, nums)).all()
    # the dense and plain forms agree on the whole slot, checksum block included
    assert (got == plain).all()


@pytest.mark.parametrize("form", FORMS)
def test_dense_config0_golden(form):
    say(form, "config0")
    c = G["config0"]
    S, n = c["slice"], c["n"]
    datas = [O.prng_bytes(0, 0, i, S) for i in range(n)]
    got, valid, _ = run_dense(datas, [S] * n, S, c["exps"], form=form)
    off, _ = P.slot_layout(S, 64)
    for r in range(len(c["exps"])):
        assert hashlib.sha256(got[r, :S].tobytes()).hexdigest() == c["sha256"][r]
        assert bytes(got[r, off:off + 64]).hex() == c["trailer"][r]
    assert (valid == 1).all()


def test_dense_rejects_bad_shapes():
    buf = P.DeviceBuffer(4096)
    assert lib().pph_muladd_dense(buf.ptr, 1000, 1, buf.ptr, buf.ptr, buf.ptr, 1, 0, None) == parpar_amd.PPH_E_INVALID
    assert lib().pph_muladd_dense(buf.ptr, 1024, 1, None, None, buf.ptr, 1, 0, None) == parpar_amd.PPH_E_INVALID


def test_bilinear_chunked(monkeypatch):
    """Several coefficient blocks and word chunks (forced small) give the same bytes."""
    say("bilinear chunked")
    n, n_out, S = 700, 75, 10000
    rng = np.random.default_rng(5)
    datas = [O.prng_bytes(72, 0, k, S) for k in range(n)]
    exps = rng.integers(0, 65535, n_out).astype(np.uint16)
    full, _, plain = run_dense(datas, [S] * n, S, exps, also_plain=True, form="bilinear")
    assert (full == plain).all()
    lib().pph_bilinear_release()
    monkeypatch.setenv("PPH_BILINEAR_ROWS", "32")
    monkeypatch.setenv("PPH_BILINEAR_CHUNK_WORDS", "1024")
    chunked, valid, _ = run_dense(datas, [S] * n, S, exps, form="bilinear")
    assert (chunked == plain).all()
    assert (valid == 1).all()
    lib().pph_bilinear_release()


@pytest.mark.parametrize("tile", ["64", "128", "256"])
@pytest.mark.parametrize("n,n_out", [(257, 300), (40, 129), (520, 256)])
def test_bilinear_row_tiles(monkeypatch, tile, n, n_out):
    """Every GEMM row tile (64 outputs x 1024 words, 128 x 512, 256 x 256)
    gives the plain form's bytes, with outputs and inputs off the tile and step
    sizes."""
    say("bilinear row tile", tile, n, n_out)
    S = 5000
    rng = np.random.default_rng(n + n_out)
    datas = [O.prng_bytes(73, n, k, S) for k in range(n)]
    exps = rng.integers(0, 65535, n_out).astype(np.uint16)
    lib().pph_bilinear_release()
    monkeypatch.setenv("PPH_BILINEAR_ROW_TILE", tile)
    got, valid, plain = run_dense(datas, [S] * n, S, exps, also_plain=True, form="bilinear")
    lib().pph_bilinear_release()
    assert (got == plain).all()
    assert (valid == 1).all()


def test_bilinear_rejects_bad_shapes():
    buf = P.DeviceBuffer(4096)
    assert lib().pph_muladd_bilinear(buf.ptr, 1000, 1, buf.ptr, buf.ptr, buf.ptr, 1, 0, None) == \
        parpar_amd.PPH_E_INVALID
    assert lib().pph_muladd_bilinear(buf.ptr, 1024, 1, None, None, buf.ptr, 1, 0, None) == parpar_amd.PPH_E_INVALID


def gf_log_table():
    """log2 over GF(2^16) with the PAR2 polynomial, from its definition."""
    log = np.zeros(65536, np.uint32)
    x = 1
    for k in range(65535):
        log[x] = k
        x <<= 1
        if x & 0x10000:
            x ^= 0x1100B
    return log


@pytest.mark.parametrize("chunked", [False, True])
def test_bilinear_explicit_coefficients(monkeypatch, chunked):
    """pph_muladd_bilinear_coef (explicit coefficients, the decoding path) equals
    the oracle's explicit-coefficient recovery, coefficients 0 and 1 included,
    also with its output block and word chunk forced small."""
    say("bilinear explicit coefficients", "chunked" if chunked else "")
    n, n_out, S, stride = 700, 75, 10000, 80
    rng = np.random.default_rng(11)
    datas = np.stack([O.prng_bytes(74, 0, k, S) for k in range(n)])
    coeffs = rng.integers(0, 65536, (n, n_out)).astype(np.uint16)
    coeffs[rng.random((n, n_out)) < 0.1] = 0
    coeffs[rng.random((n, n_out)) < 0.05] = 1
    ref = O.recover(datas, [S] * n, S, coeffs=coeffs)
    log = gf_log_table()
    table = np.full((n, stride), 0xFFFFFFFF, np.uint32)
    table[:, :n_out] = np.where(coeffs == 0, np.uint32(0xFFFFFFFF), 2 * log[coeffs])
    _, slot = P.slot_layout(S, 64)
    host = np.zeros((n, slot), np.uint8)
    host[:, :S] = datas
    src = P.DeviceBuffer(host.nbytes)
    src.upload(host)
    data = P.DeviceBuffer(host.nbytes)
    parpar_amd.check(lib().pph_prep_slots_from(data.ptr, src.ptr, slot, n, None, S, 64, 0, None), "prep")
    d_tab = P.DeviceBuffer(table.nbytes)
    d_tab.upload(table)
    out = P.DeviceBuffer(slot * n_out)
    out.upload(np.full(slot * n_out, 0x5A, np.uint8))
    lib().pph_bilinear_release()
    if chunked:
        monkeypatch.setenv("PPH_BILINEAR_ROWS", "32")
        monkeypatch.setenv("PPH_BILINEAR_CHUNK_WORDS", "1024")
    parpar_amd.check(lib().pph_muladd_bilinear_coef(data.ptr, slot, n, d_tab.ptr, stride, out.ptr, n_out, 0, None),
                     "bilinear_coef")
    valid = P.DeviceBuffer(4 * n_out)
    parpar_amd.check(lib().pph_verify_slots(out.ptr, slot, n_out, S, 64, valid.ptr, None), "verify")
    parpar_amd.check(lib().pph_synchronize())
    lib().pph_bilinear_release()
    got = out.download().reshape(n_out, slot)[:, :S]
    assert (got == ref).all()
    assert (valid.download(4 * n_out).view(np.int32) == 1).all()
    # a table narrower than the outputs is refused
    assert lib().pph_muladd_bilinear_coef(data.ptr, slot, n, d_tab.ptr, n_out - 1, out.ptr, n_out, 0, None) == \
        parpar_amd.PPH_E_INVALID
