"""Build ``libgstex_hip.so`` for gfx950 in-tree with hipcc (via csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")
LIB = os.path.join(_HERE, "libgstex_hip.so")


def build_library(jobs: int = 8, force: bool = False) -> str:
    """Compile every HIP source and link the shared library; returns its path."""
    if force:
        subprocess.run(["make", "-C", CSRC, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", CSRC, f"-j{jobs}"], check=True)
    return LIB


if __name__ == "__main__":
    print(build_library())
