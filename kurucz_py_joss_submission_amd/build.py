"""Compile libsynthe_hip.so (gfx950) in-tree with hipcc.

    python -m kurucz_py_joss_submission_amd.build

Flags: -ffp-contract=off keeps multiplies and adds separately rounded so the
kernels reproduce the reference's floating-point operation order; packed
FP32 instructions are disabled; denormals are preserved (hipcc default).
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
INCLUDE = ROOT / "include"
LIB = PKG / "libsynthe_hip.so"
ARCH = os.environ.get("SYNTHE_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose: bool = False, extra_flags=None) -> Path:
    sys.path.insert(0, str(CSRC))
    try:
        import gen_tables  # noqa: E402
    finally:
        sys.path.pop(0)
    gen_tables.main()
    srcs = [CSRC / "synthe_hip.hip"]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *COMMON_FLAGS, *(extra_flags or []), f"-I{INCLUDE}", f"-I{CSRC}",
           *map(str, srcs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
    if verbose and res.stderr.strip():
        print(res.stderr[-4000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
