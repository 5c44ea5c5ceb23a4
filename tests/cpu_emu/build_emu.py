"""Build libefhps_emu.so: the product's HIP sources compiled for the CPU against
hip_cpu_shim.h, with guard pages behind every device buffer (test infrastructure).

Source transforms (applied to copies under tests/cpu_emu/_build/):
  #include <hip/hip_runtime.h>         -> #include "hip_cpu_shim.h"
  extern __shared__ double X[];        -> double* X = (double*)emu_dyn_smem();
  __shared__                            -> static   (one block runs at a time)
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "ellipticforest_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libefhps_emu.so")
SOURCES = ["ef_kernels.hip", "ef_tree.hip", "ef_patch_kernels.hip", "ef_patch_large.hip", "ef_api.hip", "ef_patch.hip",
           "ef_mesh.cpp"]
HEADERS = ["ef_kernels.hpp", "ef_device.hpp", "ef_host.hpp", "ef_tree.hpp"]


def transform(text):
    text = text.replace("#include <hip/hip_runtime.h>", '#include "hip_cpu_shim.h"')
    text = re.sub(r"extern __shared__ (\w+) (\w+)\[\];", r"\1* \2 = (\1*)emu_dyn_smem();", text)
    text = text.replace("__shared__", "static")
    # occupancy hints are for the GPU compiler only
    text = re.sub(r"__attribute__\(\(amdgpu_waves_per_eu\([^;{]*?\)\)\)\s*", "", text)
    return text


def clangxx():
    for c in ("/opt/rocm/llvm/bin/clang++", "clang++"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    raise RuntimeError("clang++ not found")


def stale():
    """True when a product source is newer than the emulator library."""
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    srcs = [os.path.join(CSRC, n) for n in SOURCES + HEADERS] + [os.path.join(HERE, "hip_cpu_shim.cpp"),
                                                                 os.path.join(HERE, "hip_cpu_shim.h")]
    return any(os.path.getmtime(s) > t for s in srcs)


def build():
    os.makedirs(OUT, exist_ok=True)
    for name in SOURCES + HEADERS:
        with open(os.path.join(CSRC, name)) as f:
            t = transform(f.read())
        dst = os.path.join(OUT, name.replace(".hip", ".cpp"))
        with open(dst, "w") as f:
            f.write(t)
    srcs = [os.path.join(OUT, n.replace(".hip", ".cpp")) for n in SOURCES] + [os.path.join(HERE, "hip_cpu_shim.cpp")]
    cmd = [clangxx(), "-std=c++20", os.environ.get("EF_EMU_OPT", "-O1"), "-g"] + os.environ.get("EF_EMU_SAN", "").split() + [ "-fPIC", "-shared", "-pthread", "-I", OUT, "-I", HERE,
           "-I", os.path.join(ROOT, "include"), "-Wno-unused-result", "-o", LIB] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("emulator build failed")
    return LIB


if __name__ == "__main__":
    print(build())
