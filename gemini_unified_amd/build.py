"""Build libgemini_fk.so (HIP, gfx950) in-tree with hipcc.

Each translation unit is compiled to an object of its own in parallel (the two
step-engine variants take most of the time), then the objects are linked into
the shared library."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgemini_fk.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["gm_exact.hip", "gm_fast.hip", "gm_mlp.hip", "gm_synth.hip", "gm_capi.cpp", "gm_noalloc.hip"]
DEPS = SOURCES + ["gm_common.h", "gm_ctx.h", "gm_mlp.h", "gm_device.inc", "gm_kernels.inc", "gm_launch.inc",
                  "gm_persist.inc"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GPU_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(force: bool = True, verbose: bool = False) -> str:
    srcs = _sources()
    if not force and os.path.exists(LIB):
        newest = max(os.path.getmtime(os.path.join(CSRC, f)) for f in DEPS if os.path.exists(os.path.join(CSRC, f)))
        newest = max(newest, os.path.getmtime(os.path.join(HERE, "..", "include", "gemini_fk.h")))
        if os.path.getmtime(LIB) >= newest:
            return LIB
    os.makedirs(OBJDIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *GPU_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        r = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}")
        return obj

    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


# GFK file library: host C++ over the HDF5 C library (no GPU code)
FILE_LIB = os.path.join(HERE, "libgemini_gfkfile.so")
HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")


def hdf5_available() -> bool:
    return os.path.exists(os.path.join(HDF5_DIR, "include", "hdf5.h")) and \
        os.path.exists(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))


def build_file_lib(verbose: bool = False):
    """libgemini_gfkfile.so, or None when the HDF5 C library is not installed."""
    if not hdf5_available():
        return None
    inc, lib = os.path.join(HDF5_DIR, "include"), os.path.join(HDF5_DIR, "lib")
    tmp = FILE_LIB + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           os.path.join(CSRC, "gm_gfkfile.cpp"), f"-I{inc}", f"-L{lib}", "-lhdf5", f"-Wl,-rpath,{lib}", "-o", tmp]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(tmp, FILE_LIB)
    return FILE_LIB


if __name__ == "__main__":
    print(build(force=True, verbose=True))
    print(build_file_lib(verbose=True))
