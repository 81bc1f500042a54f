"""The library's host threading under ThreadSanitizer (CPU only): the task
queue of the backend's completion worker and of the Node binding's worker,
the parallel copy pool of the backend's staging and getOutput copies, and the
serving engine's host MD5 pool order.  The
check program (tests/cpp/tsan_queues.cpp) is built with g++ -fsanitize=thread
into a temporary directory and must print "ok" with no ThreadSanitizer
report."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_queues_and_copy_pool_under_tsan():
    src = os.path.join(ROOT, "tests", "cpp", "tsan_queues.cpp")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "tsan_queues")
        b = subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread", src, "-o", exe],
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
        assert b.returncode == 0, b.stdout
        env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1")
        r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300, env=env)
    assert "ThreadSanitizer" not in r.stdout, r.stdout[-4000:]
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-4000:]
