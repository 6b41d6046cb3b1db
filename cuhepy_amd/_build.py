"""Offline build of the gfx950 library ``cuhepy_amd/libchx.so`` with hipcc.

The library is one HIP translation unit per group of kernels; the units are
compiled in parallel and linked into a single shared object. No network, no
cmake: this replaces the reference's Docker + nvcc build
(``build_gpu_binaries.sh:86-147``, ``_gpu_ext/Makefile:15-48``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libchx.so"
ARCH = os.environ.get("CHX_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are kept out of every code object (no floating point
# is used on the hot path; the flag guards helper code the compiler may emit).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    # the exponentiation kernels fully unroll K rows x K positions
    "-mllvm", "-pragma-unroll-threshold=1000000",
]

# translation units; each compiles the API object or a slice of kernels
UNITS = ["chx.hip"]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found (ROCm toolchain required to build)")
    return exe


def build(verbose: bool = True, jobs: int = 8) -> Path:
    """Compile every unit for gfx950 and link libchx.so (always rebuilds)."""
    cc = hipcc()
    with tempfile.TemporaryDirectory(prefix="chx_build_") as tmp:
        objs = []

        def compile_unit(name: str) -> str:
            src = CSRC / name
            obj = os.path.join(tmp, name + ".o")
            cmd = [cc, *HIP_FLAGS, "-c", str(src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed for {name}:\n{r.stderr[-6000:]}")
            if verbose:
                noisy = [l for l in r.stderr.splitlines() if l.strip() and "packed-fp32-ops" not in l]
                if noisy:
                    print("\n".join(noisy[-40:]), file=sys.stderr)
            return obj

        with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(UNITS)))) as ex:
            objs = list(ex.map(compile_unit, UNITS))
        out_tmp = str(LIB) + ".tmp"
        cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out_tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
        os.replace(out_tmp, LIB)
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB) for {ARCH}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
