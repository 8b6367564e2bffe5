"""Builds the C++ facade driver (facade_run.cpp) against libluma_amd.so with g++, into
tests/cpp/_build/.  Called by __graft_entry__.build() after the library."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
OUT = HERE / "_build" / "facade_run"


def build_driver(verbose: bool = False) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "include"), str(HERE / "facade_run.cpp"),
           "-L", str(ROOT / "luma_amd" / "lib"), "-lluma_amd", "-Wl,-rpath,$ORIGIN/../../../luma_amd/lib",
           "-o", str(OUT)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return OUT


if __name__ == "__main__":
    print(build_driver(verbose="-v" in sys.argv))
