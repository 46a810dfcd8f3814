"""Build recipe of libfesomtra.so (gfx950 HIP kernels + the C ABI of include/fesomtra.h).

``python -m fesom2_tracer_dwarf_amd._build`` compiles in place; the shared object lands next to
this file and travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
REPO = PKG.parent
LIB = PKG / "libfesomtra.so"
SOURCES = ["fesomtra_kernels.hip", "fesomtra_abi.cpp"]
HEADERS = ["fesomtra_types.hpp", "fesomtra_column.hpp", "fesomtra_kernels.hpp"]

# -ffp-contract=off: the FP64 path keeps every a*b+c unfused, as the reference's CPU build does,
# so results are bit-identical to it.  Packed FP32 (v_pk_*_f32) instructions are disabled: some of
# their forms have produced silently wrong results on this platform.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-x", "hip",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPFLAGS, f"-I{REPO / 'include'}", f"-I{CSRC}", *srcs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libfesomtra.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
