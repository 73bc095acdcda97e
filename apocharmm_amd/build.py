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

SOURCES = ["system.hip", "util.hip", "nblist.hip", "pair.hip", "bonded.hip", "pme.hip", "md.hip", "capi.cpp"]

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
    # the C++ ForceManager mirror must compile against the C ABI header
    probe = os.path.join(OBJ, "fm_header_check.cpp")
    with open(probe, "w") as f:
        f.write('#include "apocharmm_amd/ForceManager.hpp"\n'
                'int fm_header_check() { apocharmm_amd::Topology t; '
                'apocharmm_amd::ForceManager fm(t); fm.setKappa(0.34); return fm.getForceStride(); }\n')
    r = subprocess.run([HIPCC] + FLAGS + ["-x", "hip", "-fsyntax-only", probe], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("ForceManager.hpp does not compile:\n%s" % r.stderr[-4000:])
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
