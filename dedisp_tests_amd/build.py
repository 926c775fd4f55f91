"""Builds dedisp_tests_amd/libdedisp.so in-tree for gfx950 (and the bench's
peak-rate probes, libdd_peaks.so).

Host sources (*.cpp) are compiled with g++ against the HIP runtime headers;
kernel sources (*.hip) with hipcc --offload-arch=gfx950. Objects go to
dedisp_tests_amd/_build/ (not tracked); the shared library lands next to this
file so it travels with the source tree.

    python -m dedisp_tests_amd.build        # or: python dedisp_tests_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libdedisp.so"
PEAKS = PKG / "peaks"
PEAKS_LIB = PKG / "libdd_peaks.so"
CLI = PKG / "cli"
BIN = PKG / "bin"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("DEDISP_OFFLOAD_ARCH", "gfx950")

HOST_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", f"-I{REPO / 'include'}", f"-I{CSRC}",
]
# Plan core: the reference's float/double operation order must be kept
# exactly, so no contraction into fused multiply-adds.
STRICT_FP = ["-ffp-contract=off"]
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", f"-I{REPO / 'include'}", f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *DEVICE_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        flags = HOST_FLAGS + (STRICT_FP if src.name == "plan.cpp" else [])
        cmd = ["g++", *flags, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> Path:
    """Compiles every source and links libdedisp.so. Always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.hip"))
    workers = min(len(sources), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC",
          *map(str, objs), "-o", str(tmp), f"-L{ROCM / 'lib'}", "-lamdhip64",
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    # the bench's peak-rate probes: a library of their own, not linked into libdedisp.so
    psrc = sorted(PEAKS.glob("*.hip"))
    pobjs = [_compile(src) for src in psrc]
    ptmp = PEAKS_LIB.with_suffix(".so.tmp")
    _run([str(ROCM / "bin" / "hipcc"), f"--offload-arch={ARCH}", "-shared", "-fPIC",
          *map(str, pobjs), "-o", str(ptmp), f"-L{ROCM / 'lib'}", "-lamdhip64",
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(ptmp, PEAKS_LIB)
    # command-line front ends (host only, linked against the library)
    BIN.mkdir(exist_ok=True)
    for src in sorted(CLI.glob("*.cpp")):
        exe = BIN / src.stem
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{REPO / 'include'}",
              f"-I{ROCM / 'include'}", str(src), "-o", str(exe) + ".tmp", f"-L{PKG}", "-ldedisp",
              f"-L{ROCM / 'lib'}", "-lamdhip64", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM / 'lib'}"])
        os.replace(str(exe) + ".tmp", exe)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
