"""Build the gfx950 shared library ``sw4lite_amd/libsw4hip.so`` in-tree.

Every HIP and host source is compiled by ``hipcc`` for
``--offload-arch=gfx950`` into an object of its own (in parallel, under
``sw4lite_amd/_build/``), then the objects are linked into the library.  Packed FP32 instructions are disabled
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
    "noalloc.hip",
    "timeseries.cpp",
    "deck.cpp",
    "topo.cpp",
    "stack.cpp",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    return res.returncode


def build(verbose=False, jobs=None):
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", CSRC, "-I", os.path.join(os.path.dirname(HERE), "include")]
    cmds, objs = [], []
    for s in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        cmds.append([hipcc()] + FLAGS + inc + ["-c", "-o", obj, os.path.join(CSRC, s)])
    jobs = jobs or min(len(cmds), max(1, os.cpu_count() or 1), 16)
    with ThreadPoolExecutor(jobs) as pool:
        rcs = list(pool.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed building libsw4hip.so")
    tmp = LIB + ".tmp"
    if _run([hipcc()] + FLAGS + ["-shared", "-pthread", "-o", tmp] + objs, verbose):
        raise RuntimeError("hipcc failed linking libsw4hip.so")
    os.replace(tmp, LIB)
    return LIB


def build_option_a(verbose=False):
    """Compile the kernel-level drop-in test program tests/option_a/option_a
    (C++ against include/sw4hip.h, linked to libsw4hip.so)."""
    root = os.path.dirname(HERE)
    src = os.path.join(root, "tests", "option_a", "option_a.cpp")
    exe = os.path.join(root, "tests", "option_a", "option_a")
    cmd = [hipcc(), "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", os.path.join(root, "include"), src, "-o",
           exe + ".tmp", "-L", HERE, "-lsw4hip", "-Wl,-rpath,$ORIGIN/../../sw4lite_amd"]
    if _run(cmd, verbose):
        raise RuntimeError("hipcc failed building tests/option_a/option_a")
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_option_a(verbose=True))
