"""Build recipe of the MI355X library: compiles the gfx950 HIP sources in-tree.

``python -m ellipticforest_amd.build`` (or ``__graft_entry__.build()``) always
recompiles ``ellipticforest_amd/libefhps.so`` from ``ellipticforest_amd/csrc``.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libefhps.so")

SOURCES = ["ef_kernels.hip", "ef_tree.hip", "ef_patch_kernels.hip", "ef_patch_large.hip", "ef_api.hip", "ef_patch.hip",
           "ef_mesh.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # packed FP32 instructions stay off: plain float code is compiled to scalar VALU ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel (hipcc, gfx950), then link the library."""
    import concurrent.futures
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]

    def compile_one(name):
        src = os.path.join(CSRC, name)
        obj = os.path.join(objdir, name + ".o")
        cmd = [hipcc()] + cflags + inc + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {name}:\n{r.stdout}{r.stderr}")
        return obj

    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950"] + objs + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libefhps.so failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


EXAMPLES = ["elliptic_single", "fv_solver", "thermal", "vtk_mesh"]


def build_examples(lib_dir=None, out_dir=None, verbose=False):
    """Drivers written against the drop-in C++ headers, linked to libefhps.so."""
    lib_dir = lib_dir or HERE
    out_dir = out_dir or os.path.join(HERE, "examples")
    outs = []
    for name in EXAMPLES:
        src = os.path.join(HERE, "examples", name + ".cpp")
        exe = os.path.join(out_dir, name)
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", exe, "-L", lib_dir,
               "-lefhps", "-Wl,-rpath," + lib_dir]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"building example {name} failed")
        outs.append(exe)
    return outs


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_examples(verbose=True))
