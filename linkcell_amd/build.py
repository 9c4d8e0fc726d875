"""Build the in-tree HIP library for gfx950 (MI355X).

``python -m linkcell_amd.build`` compiles ``linkcell_amd/csrc/*.hip`` into
``linkcell_amd/liblinkcell_amd.so``. The library exports the C waist
declared in ``include/*.h`` and is what every Python entry point loads.
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
LIB = PKG / "liblinkcell_amd.so"

SOURCES = ["lc_engine.hip", "lc_sched.hip", "lc_api.hip"]

# -ffp-contract=off: no FMA contraction anywhere, so every double
# operation rounds exactly as the reference CPU walk's.
# -packed-fp32-ops: keep the compiler away from packed FP32 forms.
HIPFLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(map(str, cmd))}")
    return res


def build_library(out: Path = LIB) -> Path:
    objs = []
    bdir = PKG / "_build"
    bdir.mkdir(exist_ok=True)
    for src in SOURCES:
        obj = bdir / (Path(src).stem + ".o")
        _run([hipcc(), *HIPFLAGS, "-I", str(ROOT / "include"), "-c",
              str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs,
          "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "liblcoracle.so"


if __name__ == "__main__":
    print(build_library())
    print(build_oracle())
