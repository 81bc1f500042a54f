"""Device-layer parity: synthetic fill, input preparation (gf16_cksum, log
domain), plain mul-add and checksum verification against the oracle and the
reference's golden vectors.  Calls go through the C ABI."""
import ctypes
import hashlib
import sys

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P
from golden_io import manifest

pytestmark = pytest.mark.gpu
G = manifest()
L = None


def lib():
    global L
    if L is None:
        L = parpar_amd.lib()
    return L


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_slots(datas, slice_size, ck):
    """Upload host slices (each <= slice_size bytes) into a slot array."""
    _, slot = P.slot_layout(slice_size, ck)
    buf = P.DeviceBuffer(slot * max(len(datas), 1))
    host = np.zeros((max(len(datas), 1), slot), np.uint8)
    # garbage beyond each input's length must be ignored by the preparation
    host[:] = 0xA5
    for k, d in enumerate(datas):
        host[k, :len(d)] = d
    buf.upload(host)
    return buf, slot


def prep(buf, slot, count, lens, slice_size, ck, to_log):
    d_lens = P.DeviceBuffer(4 * max(count, 1))
    d_lens.upload(np.array(lens, np.uint32))
    parpar_amd.check(lib().pph_prep_slots(buf.ptr, slot, count, d_lens.ptr, slice_size, ck, int(to_log), None),
                     "pph_prep_slots")
    parpar_amd.check(lib().pph_synchronize())
    return d_lens


def test_fill_random_matches_stream():
    say("fill_random")
    for S, ln in ((4096, 4096), (1048576, 1048576), (5000, 3333), (1024, 1)):
        _, slot = P.slot_layout(S, 64)
        buf = P.DeviceBuffer(slot * 3)
        parpar_amd.check(lib().pph_fill_random(buf.ptr, slot, 3, ln, 7, 2, 40, None), "fill")
        host = buf.download().reshape(3, slot)
        for k in range(3):
            assert (host[k, :ln] == O.prng_bytes(7, 2, 40 + k, ln)).all()
            assert not host[k, ln:].any()


@pytest.mark.parametrize("ck", [8, 16, 32, 64])
def test_prep_checksum_matches_reference(ck):
    say("prep ck", ck)
    recs = [r for r in G["cksum"] if r["width"] == ck]
    for rec in recs:
        S, src = rec["slice_len"], rec["src_len"]
        data = O.prng_bytes(*rec["key"], src) if src else np.zeros(0, np.uint8)
        for to_log in (0, 1):
            buf, slot = make_slots([data], S, ck)
            prep(buf, slot, 1, [src], S, ck, to_log)
            host = buf.download()
            off, _ = P.slot_layout(S, ck)
            trailer = host[off:off + ck]
            words = host.view(np.uint16)
            if to_log:
                # log domain: word x -> log2(x), 0 -> 0xFFFF
                lg = lambda x: 0xFFFF if x == 0 else O.lib().po_log(int(x))
                trailer = np.array([O.lib().po_exp(int(v)) if v != 0xFFFF else 0
                                    for v in trailer.view(np.uint16)], np.uint16).view(np.uint8)
                padded = np.zeros(slot, np.uint8)
                padded[:src] = data
                ref = padded.view(np.uint16)
                sample = list(range(0, min(64, slot // 2))) + list(range(max(0, S // 2 - 8), min(slot // 2, S // 2 + 8)))
                for w in sample:
                    if off // 2 <= w < (off + ck) // 2:
                        continue
                    assert words[w] == lg(ref[w]), (w, S, src)
            else:
                assert (host[:src] == data).all()
                assert not host[src:off].any()
            assert bytes(trailer).hex() == rec["trailer"], (S, src, to_log)


def test_verify_detects_corruption():
    say("verify")
    S, ck = 70000, 64
    datas = [O.prng_bytes(3, 0, k, S) for k in range(4)]
    buf, slot = make_slots(datas, S, ck)
    prep(buf, slot, 4, [S] * 4, S, ck, 0)
    valid = P.DeviceBuffer(16)
    parpar_amd.check(lib().pph_verify_slots(buf.ptr, slot, 4, S, ck, valid.ptr, None))
    assert (valid.download(16).view(np.int32) == 1).all()
    host = buf.download().reshape(4, slot)
    host[2, 12345] ^= 1
    host[3, S - 1] ^= 0x80
    buf.upload(host)
    parpar_amd.check(lib().pph_verify_slots(buf.ptr, slot, 4, S, ck, valid.ptr, None))
    assert list(valid.download(16).view(np.int32)) == [1, 1, 0, 0]


def run_plain(datas, lens, S, exps=None, nums=None, coeffs=None, group=0, ck=64, accumulate_split=None):
    """Prep inputs to log domain and run the plain kernel; returns (outputs, valid)."""
    n = len(datas)
    buf, slot = make_slots(datas, S, ck)
    prep(buf, slot, n, lens, S, ck, 1)
    n_out = len(exps) if coeffs is None else coeffs.shape[1]
    out = P.DeviceBuffer(slot * n_out)
    out.upload(np.full(slot * n_out, 0x5A, np.uint8))  # overwritten by the first batch
    d_logs = d_exps = d_coef = None
    stride = 0
    if coeffs is None:
        nums = list(range(n)) if nums is None else nums
        d_logs = P.DeviceBuffer(2 * max(n, 1))
        d_logs.upload(np.array([lib().pph_input_log(x) for x in nums] or [0], np.uint16))
        d_exps = P.DeviceBuffer(2 * n_out)
        d_exps.upload(np.array(exps, np.uint16))
    else:
        stride = (n_out + 15) // 16 * 16
        tab = np.full((max(n, 1), stride), 0xFFFFFFFF, np.uint32)
        for i in range(n):
            for r in range(n_out):
                c = int(coeffs[i, r])
                if c:
                    tab[i, r] = 2 * O.lib().po_log(c)
        d_coef = P.DeviceBuffer(tab.nbytes)
        d_coef.upload(tab)
    splits = [(0, n)] if accumulate_split is None else [(0, accumulate_split), (accumulate_split, n)]
    for bi, (a, b) in enumerate(splits):
        in_ptr = buf.offset(a * slot)
        logs_ptr = d_logs.offset(2 * a) if d_logs is not None else None
        coef_ptr = d_coef.offset(4 * stride * a) if d_coef is not None else None
        parpar_amd.check(lib().pph_muladd_plain(in_ptr, slot, b - a, logs_ptr, d_exps.ptr if d_exps else None,
                                                coef_ptr, stride, out.ptr, n_out, 1 if bi else 0, group, None),
                         "pph_muladd_plain")
    valid = P.DeviceBuffer(4 * n_out)
    parpar_amd.check(lib().pph_verify_slots(out.ptr, slot, n_out, S, ck, valid.ptr, None))
    parpar_amd.check(lib().pph_synchronize())
    host = out.download().reshape(n_out, slot)
    return host[:, :S].copy(), valid.download(4 * n_out).view(np.int32).copy()


@pytest.mark.parametrize("group", [0, 1, 2, 4, 8, 16])
def test_plain_random_vs_oracle(group):
    say("plain group", group)
    rng = np.random.default_rng(group)
    for trial in range(3):
        n = int(rng.integers(1, 40))
        S = int(rng.integers(1, 20000)) * 2
        n_out = int(rng.integers(1, 35))
        lens = [S] * n
        lens[-1] = int(rng.integers(0, S + 1))
        datas = [O.prng_bytes(50 + trial, group, k, lens[k]) for k in range(n)]
        # sprinkle zero words, which take the log-domain zero path
        if lens[0] >= 64:
            datas[0][10:40] = 0
        exps = rng.integers(0, 65535, n_out).astype(np.uint16)
        nums = rng.integers(0, 32768, n).astype(np.uint16)
        got, valid = run_plain(datas, lens, S, exps=exps, nums=list(nums), group=group,
                               accumulate_split=n // 2 if n > 1 else None)
        inp = np.zeros((n, S), np.uint8)
        for k in range(n):
            inp[k, :lens[k]] = datas[k]
        ref = O.recover(inp, lens, S, exps, input_nums=nums)
        assert (got == ref).all(), (trial, n, S, n_out)
        assert (valid == 1).all()


def test_plain_no_inputs():
    """A launch over no inputs: the sum of nothing, so overwriting zeroes the
    recovery slots (the work split has no units to run) and accumulating
    leaves them as they were."""
    say("plain no inputs")
    S = 20000
    _, slot = P.slot_layout(S, 64)
    n_out = 21
    out = P.DeviceBuffer(slot * n_out)
    before = np.full(slot * n_out, 0x5A, np.uint8)
    out.upload(before)
    d_exps = P.DeviceBuffer(2 * n_out)
    d_exps.upload(np.arange(n_out, dtype=np.uint16))
    parpar_amd.check(lib().pph_muladd_plain(out.ptr, slot, 0, None, d_exps.ptr, None, 0, out.ptr, n_out, 1, 0, None),
                     "pph_muladd_plain accumulate")
    parpar_amd.check(lib().pph_synchronize())
    assert (out.download() == before).all()
    parpar_amd.check(lib().pph_muladd_plain(out.ptr, slot, 0, None, d_exps.ptr, None, 0, out.ptr, n_out, 0, 0, None),
                     "pph_muladd_plain overwrite")
    parpar_amd.check(lib().pph_synchronize())
    assert not out.download().any()


def test_plain_explicit_coefficients():
    say("plain explicit")
    rng = np.random.default_rng(9)
    n, S, n_out = 11, 6000, 19
    datas = [O.prng_bytes(8, 0, k, S) for k in range(n)]
    coeffs = rng.integers(0, 65536, (n, n_out)).astype(np.uint16)
    coeffs[0, :] = 0
    coeffs[1, :5] = 1
    coeffs[:, 7] = 0
    got, valid = run_plain(datas, [S] * n, S, coeffs=coeffs)
    ref = O.recover(np.stack(datas), [S] * n, S, coeffs=coeffs)
    assert (got == ref).all()
    assert (valid == 1).all()


def test_plain_config0_golden():
    say("config0 device path")
    c = G["config0"]
    S, n = c["slice"], c["n"]
    ck = 64
    _, slot = P.slot_layout(S, ck)
    buf = P.DeviceBuffer(slot * n)
    parpar_amd.check(lib().pph_fill_random(buf.ptr, slot, n, S, 0, 0, 0, None))
    parpar_amd.check(lib().pph_prep_slots(buf.ptr, slot, n, None, S, ck, 1, None))
    n_out = len(c["exps"])
    d_logs = P.DeviceBuffer(2 * n)
    d_logs.upload(np.array([lib().pph_input_log(i) for i in range(n)], np.uint16))
    d_exps = P.DeviceBuffer(2 * n_out)
    d_exps.upload(np.array(c["exps"], np.uint16))
    out = P.DeviceBuffer(slot * n_out)
    parpar_amd.check(lib().pph_muladd_plain(buf.ptr, slot, n, d_logs.ptr, d_exps.ptr, None, 0, out.ptr, n_out, 0, 0,
                                            None))
    valid = P.DeviceBuffer(4 * n_out)
    parpar_amd.check(lib().pph_verify_slots(out.ptr, slot, n_out, S, ck, valid.ptr, None))
    parpar_amd.check(lib().pph_synchronize())
    host = out.download().reshape(n_out, slot)
    off, _ = P.slot_layout(S, ck)
    for r in range(n_out):
        assert hashlib.sha256(host[r, :S].tobytes()).hexdigest() == c["sha256"][r]
        assert bytes(host[r, off:off + ck]).hex() == c["trailer"][r]
    assert (valid.download(4 * n_out).view(np.int32) == 1).all()


def test_invalid_shapes_are_rejected():
    say("shape checks")
    buf = P.DeviceBuffer(4096)
    # slot size not a multiple of 1024
    assert lib().pph_muladd_plain(buf.ptr, 1000, 1, None, None, None, 0, buf.ptr, 1, 0, 0, None) == \
        parpar_amd.PPH_E_INVALID
    # unsupported group
    assert lib().pph_muladd_plain(buf.ptr, 1024, 1, buf.ptr, buf.ptr, None, 0, buf.ptr, 1, 0, 3, None) == \
        parpar_amd.PPH_E_INVALID
    # explicit table narrower than the output groups
    assert lib().pph_muladd_plain(buf.ptr, 1024, 1, None, None, buf.ptr, 4, buf.ptr, 20, 0, 16, None) == \
        parpar_amd.PPH_E_INVALID
    # slot too small for slice + checksum
    assert lib().pph_prep_slots(buf.ptr, 1024, 1, None, 1000, 64, 1, None) == parpar_amd.PPH_E_INVALID
