"""Compile the HIP sources into the in-tree shared library ``jaxdd_amd/libddhip.so`` for gfx950.

``hipcc`` cross-compiles without a GPU. Packed FP32 instructions are switched off
(``-packed-fp32-ops``): every FP32 operation in the kernels is scalar.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libddhip.so"
SOURCES = ["ddhip_solve.hip"]
ARCH = os.environ.get("DDHIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def command(out: Path = LIB) -> list[str]:
    return [
        HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(HERE.parent / "include"),
        "-o", str(out), *[str(CSRC / s) for s in SOURCES],
    ]


def build(verbose: bool = False) -> Path:
    """Compile (always) and return the library's path. The output is written to a temporary
    name first and renamed, so a process that has the old library open keeps a whole file."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
