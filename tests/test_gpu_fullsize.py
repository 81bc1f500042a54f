"""Full-size parity through properties that hold at any size, on the
benchmark workloads: 2000 inputs x 1 MiB with 200 recovery slices (plain
form) and 32768 inputs x 64 KiB with 8192 recovery slices (the bilinear form
and the int8-MFMA bit-matrix form).

- Every recovery slice's gf16_cksum checks against its data, and every
  checksum block equals the oracle's recovery run over the inputs' checksum
  blocks as if they were slices (the checksum is linear: ck(sum c_i x_i) =
  sum c_i ck(x_i)), so all R rows are pinned to the CPU, not a sample.
- Encode -> erase -> decode: erase k inputs, take k recovery slices, remove
  the surviving inputs' share (the plain kernel with explicit coefficients,
  accumulating), invert the k x k coefficient matrix over GF(2^16) and apply
  it.  The erased inputs must come back byte for byte.

The field arithmetic and the PAR2 coefficients here come from their
definitions (the 0x1100B field, input constants 2^n for n not divisible by
3, 5, 17 or 257, c(r, i) = (2^n_i)^e_r), not from the library."""
import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P

pytestmark = pytest.mark.gpu

ORDER = 65535
EXP = [0] * ORDER
LOG = [0] * 65536
_v = 1
for _k in range(ORDER):
    EXP[_k] = _v
    LOG[_v] = _k
    _v <<= 1
    if _v & 0x10000:
        _v ^= 0x1100B


def gmul(a, b):
    return 0 if a == 0 or b == 0 else EXP[(LOG[a] + LOG[b]) % ORDER]


def ginv(a):
    return EXP[(ORDER - LOG[a]) % ORDER]


def input_logs(n):
    out, e = [], 0
    while len(out) < n:
        e += 1
        if e % 3 and e % 5 and e % 17 and e % 257:
            out.append(e)
    return out


def coef_log(in_log, exp):
    return (in_log * exp) % ORDER


def invert(m):
    k = len(m)
    a = [row[:] + [1 if i == j else 0 for j in range(k)] for i, row in enumerate(m)]
    for c in range(k):
        p = next(r for r in range(c, k) if a[r][c])
        a[c], a[p] = a[p], a[c]
        inv = ginv(a[c][c])
        a[c] = [gmul(x, inv) for x in a[c]]
        for r in range(k):
            if r != c and a[r][c]:
                f = a[r][c]
                a[r] = [x ^ gmul(f, y) for x, y in zip(a[r], a[c])]
    return [row[k:] for row in a]


def say(*a):
    print(*a, flush=True)


_CK_CACHE = {}  # the oracle's checksum-block recovery per workload, shared by the forms


def roundtrip(n, S, R, form, erased, rows, seed):
    L = parpar_amd.lib()
    chk = parpar_amd.check
    ck_off, slot = P.slot_layout(S, 64)
    logs = input_logs(n)
    assert logs[:n] == [L.pph_input_log(i) for i in range(n)]
    src, prepd, out = P.DeviceBuffer(slot * n), P.DeviceBuffer(slot * n), P.DeviceBuffer(slot * R)
    d_logs, d_exps = P.DeviceBuffer(2 * n), P.DeviceBuffer(2 * R)
    d_logs.upload(np.array(logs, np.uint16))
    d_exps.upload(np.arange(R, dtype=np.uint16))
    chk(L.pph_fill_random(src.ptr, slot, n, S, seed, 0, 0, None), "fill")
    bl = form in ("bilinear", "dense")  # data-form inputs
    chk(L.pph_prep_slots_from(prepd.ptr, src.ptr, slot, n, None, S, 64, 0 if bl else 1, None), "prep")
    if form == "bilinear":
        chk(L.pph_muladd_bilinear(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, out.ptr, R, 0, None), "bilinear")
    elif form == "dense":
        chk(L.pph_muladd_dense(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, out.ptr, R, 0, None), "dense")
    else:
        chk(L.pph_muladd_plain(prepd.ptr, slot, n, d_logs.ptr, d_exps.ptr, None, 0, out.ptr, R, 0, 0, None), "plain")
    valid = P.DeviceBuffer(4 * R)
    chk(L.pph_verify_slots(out.ptr, slot, R, S, 64, valid.ptr, None), "verify")
    chk(L.pph_synchronize())
    flags = valid.download(4 * R).view(np.int32)
    assert (flags == 1).all(), np.flatnonzero(flags != 1)[:10]
    say(" encoded, every checksum valid:", n, S, R, form)
    key = (n, S, R, seed)
    if key not in _CK_CACHE:
        cks = np.empty((n, 64), np.uint8)
        for i in range(n):
            cks[i] = O.cksum(O.prng_bytes(seed, 0, i, S), S, 64)
        _CK_CACHE[key] = O.recover(cks, [64] * n, 64, np.arange(R, dtype=np.uint16))
    want_ck = _CK_CACHE[key]
    got_ck = out.download(slot * R).reshape(R, slot)[:, ck_off:ck_off + 64]
    bad = np.flatnonzero((got_ck != want_ck).any(axis=1))
    assert bad.size == 0, bad[:10]
    say(" every checksum block equals the oracle's (linearity)")

    k = len(erased)
    syn = P.DeviceBuffer(slot * k)
    for j, r in enumerate(rows):
        chk(L.pph_copy(syn.offset(j * slot), out.offset(r * slot), slot, None), "copy")
    if bl:  # the syndrome step runs the plain kernel: log-domain inputs
        chk(L.pph_prep_slots_from(prepd.ptr, src.ptr, slot, n, None, S, 64, 1, None), "prep log")
    gone = set(erased)
    kst = 16  # table stride: covers the kernel's output group (0xFFFFFFFF: no product)
    tab = np.full((n, kst), 0xFFFFFFFF, np.uint32)
    for i in range(n):
        if i not in gone:
            tab[i, :k] = [2 * coef_log(logs[i], r) for r in rows]
    d_tab = P.DeviceBuffer(tab.nbytes)
    d_tab.upload(tab)
    # syn_j = out_rows[j] + sum over surviving i of c(rows[j], i) in_i = sum over erased e of c(rows[j], e) in_e
    chk(L.pph_muladd_plain(prepd.ptr, slot, n, None, None, d_tab.ptr, kst, syn.ptr, k, 1, 0, None), "syndromes")
    chk(L.pph_prep_slots(syn.ptr, slot, k, None, S, 64, 1, None), "prep syndromes")
    m = [[EXP[coef_log(logs[e], r)] for e in erased] for r in rows]
    mi = invert(m)
    tab2 = np.full((k, kst), 0xFFFFFFFF, np.uint32)
    for j in range(k):
        for e in range(k):
            if mi[e][j]:
                tab2[j, e] = 2 * LOG[mi[e][j]]
    d_tab2 = P.DeviceBuffer(tab2.nbytes)
    d_tab2.upload(tab2)
    rec = P.DeviceBuffer(slot * k)
    chk(L.pph_muladd_plain(syn.ptr, slot, k, None, None, d_tab2.ptr, kst, rec.ptr, k, 0, 0, None), "decode")
    chk(L.pph_synchronize())
    got = rec.download(slot * k).reshape(k, slot)[:, :S]
    for j, e in enumerate(erased):
        want = O.prng_bytes(seed, 0, e, S)
        assert (got[j] == want).all(), (form, e)
    say(" erased inputs", erased, "recovered from recovery slices", rows)


def test_fullsize_plain_2000x1MiB_R200():
    say("full size plain")
    roundtrip(2000, 1 << 20, 200, "plain", erased=[0, 777, 1999], rows=[0, 100, 199], seed=1)


def test_fullsize_bilinear_32768x64KiB_R8192():
    say("full size bilinear")
    roundtrip(32768, 1 << 16, 8192, "bilinear", erased=[1, 4096, 20000, 32767], rows=[0, 3, 5000, 8191], seed=3)
    parpar_amd.lib().pph_bilinear_release()


def test_fullsize_dense_32768x64KiB_R8192():
    """The int8-MFMA GF(2) bit-matrix kernel (k_muladd_dense) at the full
    dense workload: every checksum block against the oracle, and a round trip."""
    say("full size dense (int8 bit-matrix)")
    roundtrip(32768, 1 << 16, 8192, "dense", erased=[2, 9999, 32766], rows=[1, 4000, 8190], seed=3)
