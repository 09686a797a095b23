"""In-tree build of the HIP library (gfx950).

``python -m sdfibm_amd.build`` compiles ``csrc/interact.hip`` into
``sdfibm_amd/libsdfibm_hip.so`` with hipcc.  Flags:

* ``--offload-arch=gfx950``: MI355X only;
* ``-ffp-contract=off``: no fused multiply-add, so the FP64 inside tests and
  fractions round exactly like the CPU (candidate lists bit-identical);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 math.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsdfibm_hip.so")
SOURCES = ["interact.hip", "surrogate.hip"]
HEADERS = ["shapes.h", "fraction.h", "mesh_dev.h", "dynamics.h", "surrogate.h"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-8000:])
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


DEMO = os.path.join(HERE, "host", "coupling_demo")


def build_host(verbose: bool = False) -> str:
    """The C++ host example linked against the library (plain g++, no HIP headers)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(os.path.dirname(HERE), "include"),
           "-o", DEMO, os.path.join(HERE, "host", "coupling_demo.cpp"), "-L", HERE, "-lsdfibm_hip",
           "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stderr[-4000:])
    return DEMO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
