"""Build ``libffcz.so`` (HIP kernels + C ABI) for gfx950 with hipcc, in-tree.

Usage: ``python -m ffcz_amd.build`` (or ``build()`` from Python).  Objects go
to ``ffcz_amd/_build/``; the library to ``ffcz_amd/libffcz.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libffcz.so"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("FFCZ_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code (v_pk_* FP32
# forms are avoided by design); hipcc's host pass warns about the unknown
# feature, which is harmless.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-munsafe-fp-atomics",
]
SOURCES = ["fft3d.hip", "fft_fast_f32.hip", "fft_fast_f64.hip", "fft_fast_c32.hip", "fft_fast_mfma.hip", "solver.hip",
           "batch.hip", "extract.hip", "container.hip",
           "huffman.hip", "kernels.hip", "plan_cache.cpp", "capi.cpp"]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs),
           "-l:libzstd.so.1", "-Wl,-rpath,/lib/x86_64-linux-gnu"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
