"""Host-side helpers of the engine, built from their headers: the serving
engine's MD5 (RFC 1321) against hashlib, and the backend's parallel host copy.
MD5: the RFC 1321 vectors, every padding boundary near 64-byte blocks, and
input fed in odd pieces; built with g++ and with ROCm's clang (the compiler of
the engine's host code), since the step order is pinned per compiler."""
import hashlib
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


ROCM_CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module", params=[("g++", "-O2"), (ROCM_CLANG, "-O3")], ids=["gcc", "rocm-clang"])
def md5_tool(request, tmp_path_factory):
    cxx, opt = request.param
    if shutil.which(cxx) is None:
        pytest.skip(cxx + " not present")
    exe = str(tmp_path_factory.mktemp("md5") / "md5_host_check")
    subprocess.run([cxx, opt, "-std=c++17", os.path.join(HERE, "cpp", "md5_host_check.cpp"), "-o", exe],
                   check=True)
    return exe


def digest(exe, data, pieces=()):
    r = subprocess.run([exe] + [str(p) for p in pieces], input=data, capture_output=True, check=True)
    return r.stdout.decode().strip()


def test_rfc1321_vectors(md5_tool):
    # RFC 1321 appendix A.5
    for msg, want in [(b"", "d41d8cd98f00b204e9800998ecf8427e"), (b"a", "0cc175b9c0f1b6a831c399e269772661"),
                      (b"abc", "900150983cd24fb0d6963f7d28e17f72"),
                      (b"message digest", "f96b697d7cb7938d525a2f31aaf161d0"),
                      (b"12345678901234567890123456789012345678901234567890123456789012345678901234567890",
                       "57edf4a22be3c955ac49da2e2107b67a")]:
        assert digest(md5_tool, msg) == want


@pytest.mark.parametrize("n", [1, 55, 56, 57, 63, 64, 65, 119, 120, 127, 128, 129, 1000, 65536 + 17])
def test_lengths_and_pieces(md5_tool, n):
    data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    want = hashlib.md5(data).hexdigest()
    assert digest(md5_tool, data) == want
    assert digest(md5_tool, data, [1, 63, 2, 64, 130, 7]) == want


def test_backend_copy_pool(tmp_path):
    """The backend's parallel host copy (staging into pinned memory)."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not present")
    exe = str(tmp_path / "copy_pool_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", os.path.join(HERE, "cpp", "copy_pool_check.cpp"), "-o",
                    exe], check=True)
    r = subprocess.run([exe], capture_output=True, check=True, timeout=120)
    assert r.stdout.decode().strip() == "ok"
