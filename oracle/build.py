"""Builds the CPU restatement (test oracle and CPU baseline) with g++.

Output: ``oracle/_build/liboracle.so``.  The reference itself (Rust + WGSL
on wgpu/Bevy) cannot be compiled from its own sources in this image, so there
is no ``oracle/_ref`` build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT_DIR = HERE / "_build"
LIB = OUT_DIR / "liboracle.so"
CXX = os.environ.get("CXX", "g++")

# WGSL rounding: no contraction, no fast-math, SSE single precision on x86-64.
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off", "-fno-fast-math"]


def build_oracle(verbose: bool = False) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [CXX, *FLAGS, str(HERE / "strand_oracle.cpp"), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_oracle(verbose=True))
