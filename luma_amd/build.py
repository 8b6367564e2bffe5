"""Build libluma_amd.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: python -m luma_amd.build   (or luma_amd.build.build_library()).
The library is written to luma_amd/lib/libluma_amd.so; object files to luma_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libluma_amd.so"

ARCH = os.environ.get("LUMA_AMD_ARCH", "gfx950")

# Every HIP translation unit: gfx950 only, no packed-FP32 instructions.
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-function", "-Wno-unused-command-line-argument",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(PKG.parent / "include"),
]

# (source, extra flags).  No unit lets the compiler contract: the FP64 unit keeps every
# operation unfused so that its arithmetic rounds like LUMA's x86-64 build, and the shifted
# forms write each fused multiply-add they use (fmaf), so a site update rounds the same way in
# every kernel that runs it.  The beam FEM (host code in ibm.hip) keeps sin and cos as separate
# libm calls rather than a merged sincos.
UNITS = [
    ("lbm_f64.hip", ["-ffp-contract=off"]),
    ("lbm_f32.hip", ["-ffp-contract=off"]),
    ("lbm_dual.hip", ["-ffp-contract=off"]),
    ("lbm_dual32.hip", ["-ffp-contract=off"]),
    ("api.hip", ["-ffp-contract=off"]),
    ("ibm.hip", ["-ffp-contract=off", "-fno-builtin-sin", "-fno-builtin-cos"]),
    ("adjoint.hip", ["-ffp-contract=off"]),
    ("io.hip", ["-ffp-contract=off"]),
    ("geometry.hip", ["-ffp-contract=off"]),
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; ROCm is required to build libluma_amd.so")
    return exe


def build_library(verbose: bool = False) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    objs, cmds = [], []
    for src, extra in UNITS:
        obj = OBJ / (Path(src).stem + ".o")
        cmds.append([cc, "-c", str(CSRC / src), "-o", str(obj), *COMMON, *extra])
        objs.append(str(obj))
    # the translation units are independent: compile them side by side (MAX_JOBS caps it)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1)))

    def compile_one(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        for fut in [pool.submit(compile_one, c) for c in cmds]:
            fut.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build_library(verbose="-v" in sys.argv)
    print(LIB)
