"""Compile libhorayzon.so for gfx950 with hipcc (no GPU needed)."""

from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
_CSRC = os.path.join(_HERE, "csrc")


def library_path() -> str:
    return os.path.join(_HERE, "libhorayzon.so")


def build_library(force: bool = True, jobs: int = 8) -> str:
    """Run the csrc Makefile; ``force`` rebuilds every object."""
    cmd = ["make", f"-j{jobs}"]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, cwd=_CSRC, check=True)
    return library_path()
