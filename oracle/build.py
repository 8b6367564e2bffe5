"""Build the FP64 CPU oracle (test infrastructure only) into oracle/_build/liboracle.so.

gcc, -O3, no fast-math and -ffp-contract=off: each floating-point operation rounds in the
order written, as in LUMA's own x86-64 -O3 build.  sin and cos stay separate libm calls, as
the beam element writes them (FEMElement.cpp:93-95, FEMBody.cpp:320-322): gcc would otherwise
merge a sin/cos pair of one angle into sincos, which glibc may round one ulp differently.
OpenMP is enabled for the timing mode used by the CPU baseline.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "liboracle.so"
SOURCES = ["lbm_oracle.c", "ibm_oracle.c", "fem_oracle.c"]


def build_oracle(verbose: bool = False) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
           "-fno-builtin-sin", "-fno-builtin-cos",
           "-fopenmp", "-Wall", "-o", str(tmp), *[str(HERE / s) for s in SOURCES], "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_oracle(verbose="-v" in sys.argv))
