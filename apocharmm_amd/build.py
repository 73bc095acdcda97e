"""Builds libapocharmm_amd.so (HIP for gfx950) in-tree.

    python -m apocharmm_amd.build        # or: python apocharmm_amd/build.py

Every translation unit is compiled with hipcc for gfx950 only, with packed
FP32 instructions disabled, then linked into one shared library next to this
file.  No CUDA, no hipify, no second code path.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libapocharmm_amd.so")
OBJ = os.path.join(HERE, "_build")
CPP_CALLER = os.path.join(OBJ, "fm_plugin")

SOURCES = ["system.hip", "util.hip", "nblist.hip", "pair.hip", "bonded.hip", "pme.hip", "md.hip",
           "capi.cpp", "charmm_files.cpp", "capi_charmm.cpp", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, OUT)
    # a C++ caller of the library through the ForceManager mirror, with a
    # plugin force (run by the GPU tests)
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=" + ARCH, "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "fm_plugin.cpp"), "-L" + HERE, "-lapocharmm_amd",
           "-Wl,-rpath,$ORIGIN/..", "-o", CPP_CALLER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C++ caller (ForceManager.hpp) does not build:\n%s" % r.stderr[-4000:])
    if verbose:
        print("built", OUT, "and", CPP_CALLER)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
