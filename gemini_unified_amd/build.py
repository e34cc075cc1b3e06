"""Build libgemini_fk.so (HIP, gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgemini_fk.so")
SOURCES = ["gm_exact.hip", "gm_fast.hip", "gm_mlp.hip", "gm_synth.hip", "gm_capi.cpp"]
DEPS = SOURCES + ["gm_common.h", "gm_ctx.h", "gm_mlp.h", "gm_device.inc", "gm_kernels.inc", "gm_launch.inc", "gm_persist.inc"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result"]


def build(force: bool = True, verbose: bool = False) -> str:
    if not force and os.path.exists(LIB):
        newest = max(os.path.getmtime(os.path.join(CSRC, f)) for f in DEPS)
        newest = max(newest, os.path.getmtime(os.path.join(HERE, "..", "include", "gemini_fk.h")))
        if os.path.getmtime(LIB) >= newest:
            return LIB
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True, verbose=True))
