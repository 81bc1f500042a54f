"""Recovery-packet MD5 on the device and IFSC entries, against hashlib and
zlib.  A recovery packet is hashed from its recovery set ID on
(lib/par2.js:475-513): set ID, "PAR 2.0\\0RecvSlic", the exponent as a
little-endian u32, then the recovery slice; hashlib hashes the packet
assembled that way, the device hashes the pieces as they arrive."""
import hashlib
import struct
import zlib

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P

pytestmark = pytest.mark.gpu
TYPE = b"PAR 2.0\0RecvSlic"


def say(*a):
    print(*a, flush=True)


def packet(set_id, exp, data):
    """A recovery packet as lib/par2.js:505-513 writes it (MD5 field zero)."""
    body = struct.pack("<I", exp) + bytes(data)
    hdr = b"PAR2\0PKT" + struct.pack("<Q", 64 + len(body)) + bytes(16) + set_id + TYPE
    return hdr + body


def expected(set_id, exps, rows):
    # the packet MD5 covers the packet from the set ID (offset 32) to its end
    return [hashlib.md5(packet(set_id, int(e), r)[32:]).digest() for e, r in zip(exps, rows)]


def slots_of(rows, S, fill=0xA5):
    _, slot = P.slot_layout(S, 64)
    host = np.full((len(rows), slot), fill, np.uint8)  # bytes past each row must not leak in
    for k, r in enumerate(rows):
        host[k, :len(r)] = r
    buf = P.DeviceBuffer(host.nbytes)
    buf.upload(host)
    return buf, slot


@pytest.mark.parametrize("S", [4, 28, 60, 64, 92, 100, 1024 + 28, 65536 + 4])
def test_packet_md5_whole_slices(S):
    say("packet md5", S)
    rng = np.random.default_rng(S)
    n = 70  # two wavefronts
    set_id = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
    exps = rng.integers(0, 65535, n).astype(np.uint16)
    exps[:2] = [0, 65534]
    rows = [O.prng_bytes(40, S % 977, k, S) for k in range(n)]
    buf, slot = slots_of(rows, S)
    h = P.PacketMd5(n)
    h.begin(set_id, exps)
    h.update(buf, slot, S)
    got = h.final()
    want = expected(set_id, exps, rows)
    for k in range(n):
        assert bytes(got[k]) == want[k], (S, k)
    h.close()


@pytest.mark.parametrize("S,seed", [(64, 1), (200, 2), (4096 + 36, 3), (20000, 4)])
def test_packet_md5_in_pieces(S, seed):
    """Chunked mode: each piece of even length starts at its slot's first byte
    (a new pass of the backend), with empty and 2-byte pieces among them."""
    rng = np.random.default_rng(seed)
    n = 9
    set_id = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
    exps = rng.integers(0, 65535, n).astype(np.uint16)
    rows = [O.prng_bytes(41, seed, k, S) for k in range(n)]
    cuts = sorted(set([0, S, 2, S - 2] + [int(x) * 2 for x in rng.integers(0, S // 2 + 1, 5)]))
    say("packet md5 pieces", S, cuts)
    h = P.PacketMd5(n)
    h.begin(set_id, exps)
    h.update(None, 0, 0)  # an empty piece
    for a, b in zip(cuts[:-1], cuts[1:]):
        buf, slot = slots_of([r[a:b] for r in rows], max(b - a, 4), fill=rng.integers(0, 256))
        h.update(buf, slot, b - a)
        buf.close()
    got = h.final()
    want = expected(set_id, exps, rows)
    for k in range(n):
        assert bytes(got[k]) == want[k], (S, k)
    # zeros (nothing computed) and a second message on the same handle
    h.begin(set_id, exps[::-1].copy())
    h.update(None, 0, S)
    got = h.final()
    want = expected(set_id, exps[::-1], [np.zeros(S, np.uint8)] * n)
    for k in range(n):
        assert bytes(got[k]) == want[k], k
    h.close()


def test_packet_md5_rejects():
    lib = parpar_amd.lib()
    h = P.PacketMd5(2)
    buf = P.DeviceBuffer(64)
    assert lib.pph_packet_md5_update(h._h, buf.ptr, 64, 8, None) == parpar_amd.PPH_E_STATE  # before begin
    h.begin(bytes(16), [1, 2])
    assert lib.pph_packet_md5_update(h._h, buf.ptr, 64, 7, None) == parpar_amd.PPH_E_INVALID  # odd length
    assert lib.pph_packet_md5_update(h._h, buf.ptr, 64, 62, None) == parpar_amd.PPH_E_INVALID  # past the slot
    with pytest.raises(parpar_amd.PPHError):
        h.begin(bytes(16), [1, 65535])
    h.close()


@pytest.mark.parametrize("S", [4, 100, 65536 + 12])
def test_slice_hash_ifsc(S):
    say("ifsc", S)
    n = 33
    rng = np.random.default_rng(S)
    lens = rng.integers(0, S + 1, n)
    lens[0] = S
    rows = [O.prng_bytes(42, S % 991, k, int(ln)) for k, ln in enumerate(lens)]
    buf, slot = slots_of(rows, S)
    d_lens = P.DeviceBuffer(4 * n)
    d_lens.upload(np.array(lens, np.uint32))
    out = P.DeviceBuffer(20 * n)
    parpar_amd.check(parpar_amd.lib().pph_slice_hash_ifsc(buf.ptr, slot, n, d_lens.ptr, S, out.ptr, None), "ifsc")
    parpar_amd.check(parpar_amd.lib().pph_synchronize())
    ifsc = out.download(20 * n).reshape(n, 20)
    for k, r in enumerate(rows):
        padded = np.zeros(S, np.uint8)
        padded[:len(r)] = r
        assert bytes(ifsc[k, :16]) == hashlib.md5(padded.tobytes()).digest(), k
        assert bytes(ifsc[k, 16:]) == struct.pack("<I", zlib.crc32(padded.tobytes()) & 0xffffffff), k


@pytest.mark.parametrize("form", [parpar_amd.FORM_PLAIN, parpar_amd.FORM_BILINEAR])
def test_backend_packet_md5_chunked(form):
    """Two passes over a slice in chunks (setCurrentSliceSize), the packet MD5
    appended from the backend's device outputs after each pass."""
    S, T = 6000, 3998  # passes of T and S - T bytes
    n_in, exps = 7, np.array([0, 5, 200, 65534], np.uint16)
    set_id = bytes(range(16))
    inp = np.stack([O.prng_bytes(43, 0, k, S) for k in range(n_in)])
    be = P.PAR2ProcHIP(0, 2)
    be.setSliceSize(S)
    be.init(cksum_size=64, input_batch=4, form=form)
    h = P.PacketMd5(len(exps))
    h.begin(set_id, exps)
    for a, b in ((0, T), (T, S)):
        be.discardOutput()
        be.setCurrentSliceSize(b - a)
        be.setRecoverySlices(exps)
        for k in range(n_in):
            be.addInput(inp[k, a:b].copy(), b - a, inputNum=k)
        be.endInput()
        h.update_from(be)
    got = h.final()
    ref = O.recover(inp, [S] * n_in, S, exps)
    want = expected(set_id, exps, ref)
    for r in range(len(exps)):
        assert bytes(got[r]) == want[r], r
    h.close()
    be.close()
