"""Build libsarsolver_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python -m sarsolver_amd.build
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
LIB_NAME = "libsarsolver_hip.so"
LIB_PATH = PKG / LIB_NAME

SOURCES = ["sar_kernels.hip", "sar_plan.hip", "conv_kernels.hip", "resample_kernels.hip",
           "scene_kernels.hip", "fft_kernels.hip"]

# -ffp-contract=off: FP64 range arithmetic keeps the reference's roundings.
# Packed FP32 instructions are disabled for all device code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the gfx950 library)")


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link the shared library. Always rebuilds."""
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    cc = hipcc()
    objs, cmds = [], []
    for src in SOURCES:
        obj = build_dir / (Path(src).stem + ".o")
        cmds.append([cc, *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src),
                     "-o", str(obj)])
        objs.append(str(obj))
    # the sources compile independently: run them side by side
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(workers) as pool:
        list(pool.map(run, cmds))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp),
           "-L/opt/rocm/lib", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB_PATH)
