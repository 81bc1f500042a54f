"""The fp4 GEMM kernel reads its accumulators with inline-asm moves, which the
compiler's hazard tracking does not see: every read must come after the
explicit wait-state padding that follows the item's last MFMA.  Checked on the
gfx950 assembly of both kernel instantiations (built here, no GPU)."""
import os
import re
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not present")
    out = str(tmp_path_factory.mktemp("blm") / "blm.s")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-S", "--cuda-device-only", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "parpar_amd", "csrc", "pph_bl_mfma.hip"), "-o", out],
                   check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return open(out).read()


@pytest.mark.parametrize("rows", ["128", "256"])
def test_accumulator_reads_follow_padding(asm, rows):
    name = re.search(r"^(_ZN3pph10k_blm_gemmILj%sE\w+):" % rows, asm, re.M).group(1)
    body = asm[asm.index(name + ":"):asm.index(".Lfunc_end", asm.index(name + ":"))]
    ins = [l.strip() for l in body.splitlines() if l.strip() and not l.strip().startswith((";", "."))]
    reads = [k for k, l in enumerate(ins) if l.startswith("v_accvgpr_read_b32")]
    assert len(reads) == 256
    first = reads[0]
    # the four s_nop 7 (32 wait states) directly before the first read, no MFMA after them
    assert ins[first - 4:first] == ["s_nop 7"] * 4, ins[first - 6:first + 1]
    assert not any(l.startswith("v_mfma") for l in ins[first - 4:reads[-1]])
