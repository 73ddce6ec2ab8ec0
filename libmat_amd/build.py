"""Builds libmat_amd/_lib/librpd_amd.so in-tree for gfx950.

Host C++ (input preparation) is compiled with g++, the HIP kernels and the
C-ABI engine with hipcc; every object is compiled with contraction disabled
(-ffp-contract=off) so host and device binary64 arithmetic round alike.
Usage: python -m libmat_amd.build [--jobs N]
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "_lib", "librpd_amd.so")
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("RPD_OFFLOAD_ARCH", "gfx950")

HOST_SRCS = ["csrc/host/rpd_prep.cpp", "csrc/host/rpd_compat.cpp",
             "csrc/host/rpd_cellhost.cpp"]
# librpd_amd_core.so leaves out the ConvexCellHost methods and adjacency
# (rpd_cellhost.cpp) for programs that compile LibMAT's voronoi_defs.cxx
CORE_OUT = os.path.join(HERE, "_lib", "librpd_amd_core.so")
CELLHOST_OBJ = "rpd_cellhost.cpp.o"
HIP_SRCS = ["csrc/device/rpd_engine.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-fno-fast-math", f"-I{INC}"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" +
                           r.stdout + r.stderr)
    return r


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    hipcc = _hipcc()
    cmds, objs = [], []
    for src in HOST_SRCS:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append(["g++"] + COMMON + ["-Wall", "-c",
                                        os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    for src in HIP_SRCS:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append([hipcc, f"--offload-arch={ARCH}"] + COMMON + [
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-c",
            os.path.join(HERE, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, cmds):
            if verbose:
                sys.stdout.write(r.stderr)
    for out, these in ((OUT, objs),
                       (CORE_OUT, [o for o in objs
                                   if os.path.basename(o) != CELLHOST_OBJ])):
        tmp = out + ".tmp"
        _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-o", tmp] +
             these + ["-lpthread"])
        os.replace(tmp, out)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
