"""Build the native libraries in-tree.

    python -m bit_amd.build            # bit_amd/libbit_amd.so (+ libtree3d.so)

libbit_amd.so is the C-ABI of include/bit_amd.h: host tree build (g++,
-ffp-contract=off so the tree rounds like the reference), the API (hipcc host
code) and the gfx950 kernels (hipcc, packed FP32 disabled).
libtree3d.so is the drop-in for the reference's C++ API (bit_amd/src/tree3d.h),
linked against libbit_amd.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
SRC = os.path.join(PKG, "src")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libbit_amd.so")
LIBTREE = os.path.join(PKG, "libtree3d.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) is disabled for every HIP translation unit.
NO_PK = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(obj, *deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(incremental: bool = False) -> str:
    """Compile everything (incremental=True skips objects newer than their sources)."""
    os.makedirs(BUILD, exist_ok=True)
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs.append(os.path.join(ROOT, "include", "bit_amd.h"))
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    tree_o = os.path.join(BUILD, "tree_build.o")
    api_o = os.path.join(BUILD, "bit_api.o")
    ker_o = os.path.join(BUILD, "kernels.o")
    def need(obj, src):
        return not incremental or _stale(obj, src, *hdrs)

    src = os.path.join(CSRC, "tree_build.cpp")
    if need(tree_o, src):
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall",
              *inc, "-c", src, "-o", tree_o])
    # the far kernel's asm sweep is generated (far_sweep_asm.h, rewritten only if it changed)
    _run([sys.executable, os.path.join(CSRC, "gen_far_sweep.py")])
    src = os.path.join(CSRC, "kernels.hip")
    if need(ker_o, src):
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
              *NO_PK, *inc, "-c", src, "-o", ker_o])
    src = os.path.join(CSRC, "bit_api.cpp")
    if need(api_o, src):
        _run([HIPCC, "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off",
              "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", *inc, "-c", src, "-o", api_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, tree_o, ker_o, api_o,
          "-lgomp"])
    os.replace(tmp, LIB)
    if os.path.isdir(SRC) and os.path.exists(os.path.join(SRC, "tree3d.cpp")):
        tmp = LIBTREE + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
              "-Wall", "-I" + SRC, *inc, os.path.join(SRC, "tree3d.cpp"), "-o", tmp,
              "-L" + PKG, "-lbit_amd", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, LIBTREE)
        # caller of the reference interface used by tests/test_dropin.py
        drv = os.path.join(ROOT, "tests", "cpp", "dropin_treecode.cpp")
        if os.path.exists(drv):
            os.makedirs(os.path.join(ROOT, "tests", "cpp", "bin"), exist_ok=True)
            _run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I" + SRC, "-I" + os.path.join(ROOT, "include"),
                  drv, "-o", os.path.join(ROOT, "tests", "cpp", "bin", "dropin_treecode"),
                  "-L" + PKG, "-ltree3d", "-lbit_amd", "-Wl,-rpath," + PKG])
    drv = os.path.join(ROOT, "tests", "cpp", "solver_dense.cpp")
    if os.path.exists(drv):
        os.makedirs(os.path.join(ROOT, "tests", "cpp", "bin"), exist_ok=True)
        _run(["g++", "-O2", "-std=c++17", "-I" + SRC, drv, "-o",
              os.path.join(ROOT, "tests", "cpp", "bin", "solver_dense")])
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
