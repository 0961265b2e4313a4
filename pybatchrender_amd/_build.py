"""Build libpbr_hip.so in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRCS = [PKG / "csrc" / "pbr_render.hip", PKG / "csrc" / "pbr_backward.hip",
        PKG / "csrc" / "pbr_env.hip", PKG / "csrc" / "pbr_tiles.hip",
        PKG / "csrc" / "pbr_rows.hip"]
OUT = PKG / "libpbr_hip.so"

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # the bit-exact contract: no contraction of a*b+c into fma unless written
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_hip(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(PKG.parent / "include"), "-o", str(tmp), *map(str, SRCS)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(res.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_hip(verbose=True))
