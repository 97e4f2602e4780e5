"""Compile the HIP/C++ sources of ``gospl_amd/csrc`` into ``libgospl_hip.so``.

One hipcc invocation, gfx950 only, in-tree output.  FMA contraction is off so
that the kernels evaluate goSPL's floating-point expressions in the same order
as the CPU code they are checked against; packed-FP32 instruction selection is
disabled.
"""

from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgospl_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GOSPL_OFFLOAD_ARCH", "gfx950")


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def build(verbose: bool = True) -> str:
    srcs = sources()
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = [HIPCC, "--offload-arch=%s" % ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", inc, "-I", CSRC, "-o", OUT + ".tmp"] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
