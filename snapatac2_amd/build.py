"""Builds libsnapgpu.so (the HIP sources in csrc/) and the bench-only
libsnapgpu_probe.so (csrc/probe/) for gfx950, in-tree.

Usage: python -m snapatac2_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libsnapgpu.so"
ARCH = os.environ.get("SNAPGPU_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have given
    # intermittent wrong results on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


PROBE_LIB = OUT_DIR / "libsnapgpu_probe.so"


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(srcs, out: Path, includes, verbose: bool) -> None:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, *(f"-I{d}" for d in includes), *map(str, srcs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build(verbose: bool = True) -> Path:
    """libsnapgpu.so (the product) and libsnapgpu_probe.so (bench.py's
    measurement aids, csrc/probe/), both for gfx950."""
    OUT_DIR.mkdir(exist_ok=True)
    _compile(sources(), LIB, [ROOT / "include", CSRC], verbose)
    _compile(sorted((CSRC / "probe").glob("*.hip")), PROBE_LIB, [CSRC / "probe"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
