"""Device hashes of the input path: per-slice MD5 / CRC-32 of the zero-padded
slice and the whole-file MD5, against Python's hashlib / zlib (RFC 1321, IEEE
CRC-32) and the committed vectors."""
import hashlib
import zlib

import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import proc as P
from golden_io import manifest

pytestmark = pytest.mark.gpu
G = manifest()


def say(*a):
    print(*a, flush=True)


def lib():
    return parpar_amd.lib()


def device_slots(datas, S):
    _, slot = P.slot_layout(S, 64)
    host = np.full((len(datas), slot), 0xEE, np.uint8)  # garbage past each length
    for k, d in enumerate(datas):
        host[k, :len(d)] = d
    buf = P.DeviceBuffer(host.nbytes)
    buf.upload(host)
    return buf, slot


def slice_hash(datas, S):
    buf, slot = device_slots(datas, S)
    n = len(datas)
    lens = P.DeviceBuffer(4 * n)
    lens.upload(np.array([len(d) for d in datas], np.uint32))
    md5 = P.DeviceBuffer(16 * n)
    crc = P.DeviceBuffer(4 * n)
    parpar_amd.check(lib().pph_slice_hash(buf.ptr, slot, n, lens.ptr, S, md5.ptr, crc.ptr, None), "slice_hash")
    parpar_amd.check(lib().pph_synchronize())
    return md5.download(16 * n).reshape(n, 16), crc.download(4 * n).view(np.uint32)


# slice sizes span every split of the CRC over lanes (1, 2, 4, 16 and 64 lanes
# per slice, with and without a lead of virtual zero bytes)
@pytest.mark.parametrize("S", [4, 60, 64, 100, 124, 128, 4096 + 52, 8192, 20000, 65536 + 12, 262144, 1048576 + 4])
def test_slice_md5_crc(S):
    say("slice hash", S)
    rng = np.random.default_rng(S)
    lens = [S, 0, 1, S - 1, int(rng.integers(0, S + 1)), 55 if S > 55 else S, 56 if S > 56 else S, 64 if S > 64 else S]
    datas = [O.prng_bytes(31, S % 1000, k, ln) for k, ln in enumerate(lens)]
    md5, crc = slice_hash(datas, S)
    for k, d in enumerate(datas):
        padded = np.zeros(S, np.uint8)
        padded[:len(d)] = d
        assert bytes(md5[k]).hex() == hashlib.md5(padded.tobytes()).hexdigest(), (S, k, len(d))
        assert int(crc[k]) == zlib.crc32(padded.tobytes()) & 0xffffffff, (S, k, len(d))


@pytest.mark.parametrize("S,n", [(8192, 300), (65536 + 12, 40)])
def test_slice_hash_many(S, n):
    """Many slices of ragged lengths: the CRC lanes of one slice and the
    slices of one launch spread over several workgroups."""
    say("slice hash many", S, n)
    rng = np.random.default_rng(n)
    lens = rng.integers(0, S + 1, n)
    lens[::7] = S
    datas = [O.prng_bytes(33, n, k, int(ln)) for k, ln in enumerate(lens)]
    md5, crc = slice_hash(datas, S)
    for k, d in enumerate(datas):
        padded = np.zeros(S, np.uint8)
        padded[:len(d)] = d
        assert bytes(md5[k]).hex() == hashlib.md5(padded.tobytes()).hexdigest(), (S, k, len(d))
        assert int(crc[k]) == zlib.crc32(padded.tobytes()) & 0xffffffff, (S, k, len(d))


def test_slice_hash_golden_vectors():
    for rec in G["hash"]:
        S, ln = rec["slice"], rec["len"]
        data = O.prng_bytes(*rec["key"], ln) if ln else np.zeros(0, np.uint8)
        md5, crc = slice_hash([data], S)
        assert bytes(md5[0]).hex() == rec["md5"]
        assert "%08x" % int(crc[0]) == rec["crc32"]


@pytest.mark.parametrize("S,file_size", [(64, 64), (64, 1), (100, 250), (4096, 4096 * 3 - 7), (262144, 262144 * 5),
                                         (1048576 + 4, 3 * 1048576 + 11), (8, 0)])
def test_file_md5(S, file_size):
    say("file md5", S, file_size)
    n = max(1, (file_size + S - 1) // S)
    data = O.prng_bytes(32, S % 997, 0, file_size) if file_size else np.zeros(0, np.uint8)
    datas = [data[k * S:(k + 1) * S] for k in range(n)]
    buf, slot = device_slots(datas, S)
    out = P.DeviceBuffer(16)
    parpar_amd.check(lib().pph_file_md5(buf.ptr, slot, n, S, file_size, out.ptr, None), "file_md5")
    parpar_amd.check(lib().pph_synchronize())
    assert bytes(out.download(16)).hex() == hashlib.md5(data.tobytes()).hexdigest()


def test_hash_rejects_bad_shapes():
    buf = P.DeviceBuffer(4096)
    assert lib().pph_slice_hash(buf.ptr, 1024, 1, None, 102, buf.ptr, buf.ptr, None) == parpar_amd.PPH_E_INVALID
    assert lib().pph_file_md5(buf.ptr, 1024, 1, 1000, 2000, buf.ptr, None) == parpar_amd.PPH_E_INVALID
