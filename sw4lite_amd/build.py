"""Build the gfx950 shared library ``sw4lite_amd/libsw4hip.so`` in-tree.

All HIP and host sources are compiled by one ``hipcc`` invocation for
``--offload-arch=gfx950``.  Packed FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).  Run as
``python -m sw4lite_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsw4hip.so")

SOURCES = [
    "kernels_basic.hip",
    "kernels_step.hip",
    "kernels_fast.hip",
    "kernels_curv.hip",
    "kernels_comp.hip",
    "kernels_adj.hip",
    "sbp_coefficients.cpp",
    "setup.cpp",
    "ew.cpp",
    "adjoint.cpp",
    "batch.cpp",
    "abi.cpp",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + FLAGS + ["-I", CSRC, "-I", os.path.join(os.path.dirname(HERE), "include"),
                               "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libsw4hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
