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
    ins = [l.strip() for l in body.splitlines() if l.strip() and not l.strip().startswith(";")]
    # the explicit padding: four s_nop 7 (32 wait states), once in the kernel
    pads = [k for k in range(len(ins) - 3) if ins[k:k + 4] == ["s_nop 7"] * 4]
    assert len(pads) == 1, pads
    k = pads[0] + 4
    # then the 256 asm reads of the accumulators, in the same block, no MFMA among them
    # (other v_accvgpr_read in the kernel are the compiler's own copies, whose
    # wait states it inserts itself)
    reads = 0
    while k < len(ins) and not ins[k].startswith(".") and reads < 256:
        assert not ins[k].startswith("v_mfma"), ins[k]
        reads += ins[k].startswith("v_accvgpr_read_b32")
        k += 1
    assert reads == 256
