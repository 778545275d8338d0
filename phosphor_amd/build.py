"""Builds the gfx950 shared library (and, for the test suite, the CPU oracle).

The product library is compiled in-tree with hipcc into
``phosphor_amd/build/libphosphor_hip.so``.  Floating point is IEEE throughout:
no fast-math, no FP contraction, correctly rounded f32 divide/sqrt and IEEE
denormals, so kernel results are reproducible bit for bit.  Packed FP32
instructions are disabled (``-packed-fp32-ops``).

Run ``python -m phosphor_amd.build`` or call :func:`build`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "build"
LIB = BUILD / "libphosphor_hip.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "build" / "librefcpu.so"

SOURCES = ["kernels_trace.hip", "kernels_photon_map.hip", "kernels_batch.hip", "kernels_adjoint.hip", "abi.cpp",
           "renderer.cpp", "batch.cpp", "adjoint.cpp", "scene_upload.cpp", "bvh.cpp", "accel.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _run(cmd: list[str], cwd: Path) -> None:
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ... (exit {res.returncode})")


def build_library(jobs: int = 8) -> Path:
    """Compile each source to an object in parallel, then link the .so."""
    BUILD.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc, *HIPCC_FLAGS, *lang, "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append(out)
    if failed:
        sys.stderr.write("\n".join(failed))
        raise RuntimeError("hipcc compilation failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp), "-lpthread"], ROOT)
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure only)."""
    _run(["make", "-s", "-C", str(ORACLE_DIR)], ROOT)
    return ORACLE_LIB


def build() -> None:
    build_library()
    build_oracle()


if __name__ == "__main__":
    build()
    print(LIB)
