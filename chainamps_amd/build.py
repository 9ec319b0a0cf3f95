"""Build the native library in-tree: chainamps_amd/lib/libchainamps_amd.so.

Every source under chainamps_amd/csrc is compiled for gfx950 with hipcc
(host code included) and linked with rocBLAS/rocSOLVER.  Usage:

    python -m chainamps_amd.build [--jobs N] [--debug]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIBDIR = HERE / "lib"
OBJDIR = HERE / "lib" / "obj"
LIBNAME = "libchainamps_amd.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("CAM_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: no per-chunk copies between VGPRs and AGPRs
    # around the grouped GEMM's K loop
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
    f"-I{ROCM}/include",
]


def hipcc() -> str:
    p = shutil.which("hipcc") or f"{ROCM}/bin/hipcc"
    if not Path(p).exists():
        raise RuntimeError("hipcc not found; the ROCm toolchain is required")
    return p


def sources():
    return sorted(list(CSRC.glob("*.cpp")) + list(CSRC.glob("*.hip")))


def compile_one(src: Path, opt: list[str]) -> Path:
    obj = OBJDIR / (src.name + ".o")
    cmd = [hipcc()] + COMMON + opt + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 0, debug: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    # objects of deleted or renamed sources must not survive a build
    if OBJDIR.exists():
        shutil.rmtree(OBJDIR)
    OBJDIR.mkdir()
    opt = ["-O0", "-g"] if debug else ["-O3", "-DNDEBUG"]
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, opt), srcs))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", str(tmp)] + [str(o) for o in objs] + [
        f"-L{ROCM}/lib", "-lrocsolver", "-lrocblas", f"-Wl,-rpath,{ROCM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    build_cpp_examples(out)
    return out


def build_cpp_examples(lib: Path) -> list[Path]:
    """The C++ programs under tests/cpp use the header-only layer
    include/chainamps_amd.hpp (the reference-shaped API) over the library;
    they are host C++ (g++, C++11) linked against it."""
    root = HERE.parent
    srcs = sorted((root / "tests" / "cpp").glob("*.cpp"))
    gxx = shutil.which("g++")
    if not srcs or not gxx:
        return []
    outs = []
    for src in srcs:
        exe = LIBDIR / src.stem
        tmp = LIBDIR / (src.stem + ".tmp")
        cmd = [gxx, "-std=c++11", "-O2", "-Wall", f"-I{root / 'include'}", str(src), "-o", str(tmp),
               f"-L{LIBDIR}", "-lchainamps_amd", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, exe)
        outs.append(exe)
    return outs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args(argv)
    out = build(a.jobs, a.debug)
    print(out)


if __name__ == "__main__":
    sys.exit(main())
