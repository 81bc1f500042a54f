"""Build libparpar_hip.so for gfx950 in-tree (parpar_amd/lib/).

    python -m parpar_amd.build          # or __graft_entry__.build()

HIP translation units are compiled with hipcc for gfx950; the PAR2ProcHIP
host class (controller_hip.cpp) is plain C++ built with g++ against the C ABI
only.  Object files go to parpar_amd/_build/.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libparpar_hip.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("PPH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_SOURCES = ["pph_kernels.hip", "pph_dense.hip", "pph_batch.hip", "pph_hash.hip", "pph_serve.hip", "pph_bilinear.hip", "pph_bl_mfma.hip",
               "pph_api.hip"]
CXX_SOURCES = ["controller_hip.cpp"]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             # packed FP32 VALU forms stay off in every device build
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + INCLUDE, "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + CSRC, "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # translation units compile independently: a few at a time
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(verbose=False):
    """The C restatement used by tests (and, where the reference is present, its build)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "liboracle.so"])
    if os.path.isdir("/root/reference/gf16"):
        _run(["make", "-C", odir, "-j8", "ref"])
        _run(["make", "-C", odir, "integration"])
        _run(["make", "-C", odir, "-j8", "uv", "portable"])
    if verbose:
        print("built oracle")


def build_test_drivers(verbose=False):
    """C++ drivers of PAR2ProcHIP used by the GPU tests (tests/cpp/)."""
    tdir = os.path.join(ROOT, "tests", "cpp")
    if not os.path.isdir(tdir):
        return
    for name in sorted(os.listdir(tdir)):
        if not name.endswith(".cpp") or name.startswith("tsan_"):
            continue  # tsan_*: built with -fsanitize=thread by tests/test_tsan.py
        exe = os.path.join(tdir, name[:-4])
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-I" + CSRC, "-I" + os.path.join(ROOT, "oracle"),
              os.path.join(tdir, name), "-o", exe, "-L" + LIBDIR, "-lparpar_hip", "-Wl,-rpath," + LIBDIR,
              "-L" + os.path.join(ROOT, "oracle"), "-loracle", "-Wl,-rpath," + os.path.join(ROOT, "oracle"),
              "-fopenmp"])
        if verbose:
            print("built", exe)


NODE_INCLUDE = "/usr/include/node"
NODE_ADDON = os.path.join(LIBDIR, "gfproc_hip.node")


def build_node_addon(verbose=False):
    """The Node binding (GfProc over PAR2ProcHIP), where Node's headers are installed."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    tmp = NODE_ADDON + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I" + NODE_INCLUDE, "-I" + INCLUDE,
          os.path.join(HERE, "node", "gfproc_hip.cc"), "-o", tmp, "-L" + LIBDIR, "-lparpar_hip",
          "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(tmp, NODE_ADDON)
    if verbose:
        print("built", NODE_ADDON)
    return NODE_ADDON


def build_all(verbose=False):
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    build_library(verbose)
    build_node_addon(verbose)
    build_oracle(verbose)
    build_test_drivers(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
