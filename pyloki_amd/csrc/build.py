"""Build the pyloki HIP library for gfx950: ``python -m pyloki_amd.csrc.build``.

Compiles every ``*.hip`` in this directory with ``hipcc`` into
``pyloki_amd/libpylokihip.so`` (in-tree, so the built library travels with the tree).

Flags: strict IEEE arithmetic (``-ffp-contract=off``: the FP64 resolve and the exact
scorer must not contract multiply-adds), no packed FP32 instructions, gfx950 only.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
LIB = PKG / "libpylokihip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]

# Every kernel of the path is the project's own: no vendor library is linked.
LINK_FLAGS: list[str] = []


def hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(cand).exists():
        msg = "hipcc not found: the HIP library cannot be built"
        raise RuntimeError(msg)
    return cand


def build(verbose: bool = False) -> Path:
    """Compile all HIP sources into the shared library (always recompiles)."""
    sources = sorted(str(p) for p in CSRC.glob("*.hip"))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *sources, *LINK_FLAGS]
    if verbose:
        print(" ".join(cmd), flush=True)
    env = dict(os.environ)
    res = subprocess.run(cmd, capture_output=True, text=True, env=env)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        sys.stderr.write(res.stderr)
        msg = f"hipcc failed with exit code {res.returncode}"
        raise RuntimeError(msg)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
