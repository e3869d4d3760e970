"""Build libxadv.so (HIP, gfx950) in-tree with hipcc.

The library is the product's compute path; it is built here, next to the
sources, so it travels with the tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "xadv.hip")
LIB = os.path.join(HERE, "libxadv.so")
NN_SRC = os.path.join(HERE, "csrc", "nn_kernels.hip")
NN_LIB = os.path.join(HERE, "libxadv_nn.so")
ARCH = os.environ.get("XADV_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every + and * rounds separately, in source order, so the
# kernels reproduce the reference arithmetic.  Packed FP32 instructions are
# switched off for the device code.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _hipcc(src, lib, verbose, shared=True):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = lib + ".tmp"
    flags = HIPCC_FLAGS if shared else [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    cmd = [hipcc, *flags, "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed building {os.path.basename(lib)}")
    os.replace(tmp, lib)
    return lib


def build(verbose: bool = False) -> str:
    """Build the solver library and the surrogate library; returns the solver library's path."""
    _hipcc(NN_SRC, NN_LIB, verbose)
    return _hipcc(SRC, LIB, verbose)


if __name__ == "__main__":
    print(build(verbose=True))
