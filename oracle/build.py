"""Build recipe for the oracle's C++ restatement (test and benchmark
infrastructure): oracle/refmatvec.cpp -> oracle/_build/refmatvec, the
reference-order H_eff matvec timed as the CPU baseline.

    python -m oracle.build
"""
from __future__ import annotations

import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build"
REFMATVEC = OUT / "refmatvec"


def build_refmatvec() -> Path:
    cxx = shutil.which("g++")
    if cxx is None:
        raise RuntimeError("g++ not found: cannot build oracle/_build/refmatvec")
    OUT.mkdir(exist_ok=True)
    tmp = OUT / "refmatvec.tmp"
    cmd = [cxx, "-O3", "-march=x86-64-v3", "-std=c++17", "-pthread", "-Wall", "-Wextra",
           "-o", str(tmp), str(HERE / "refmatvec.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    tmp.replace(REFMATVEC)
    return REFMATVEC


if __name__ == "__main__":
    print(build_refmatvec())
    sys.exit(0)
