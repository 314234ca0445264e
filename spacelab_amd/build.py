"""Build the spacelab_amd HIP library for gfx950, in-tree.

    python -m spacelab_amd.build          # -> spacelab_amd/lib/libspacelab.so

Every source is compiled with hipcc for gfx950 only, with FMA contraction off
(the ordered kernels must issue exactly the reference's binary64 operations)
and packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libspacelab.so")

SOURCES = ["sl_api.hip", "sl_kernels_ordered.hip", "sl_kernels_fast.hip", "sl_util.hip",
           "sl_batch.hip", "sl_mlp.hip", "sl_static.hip", "sl_host.cpp", "sl_tune.cpp", "sl_collide.cpp", "sl_csv.cpp"]
HOSTDIR = os.path.join(PKG, "host")
# host programs over the C ABI: (source, executable in lib/)
PROGRAMS = [("collider_main.cpp", "spacelab_collider")]
HEADERS = ["sl_common.hpp", "sl_pair_ref.hpp", "sl_internal.hpp", "sl_fast_pair.hpp",
           "sl_fast_kernels.hpp"]

ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall", "-Wno-unused-function"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON
    if src.endswith(".hip"):
        cmd += DEVICE
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link libspacelab.so.  Always rebuilds."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for src, exe in PROGRAMS:
        out = os.path.join(LIBDIR, exe)
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
               "-Wall", "-pthread", "-I" + os.path.join(ROOT, "include"), "-I" + HOSTDIR, os.path.join(HOSTDIR, src),
               "-L" + LIBDIR, "-lspacelab", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host program build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(out + ".tmp", out)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
