"""Build the engine and QuEST API libraries in-tree (dgquest_amd/lib/).

Runs the package Makefile: hipcc for the gfx950 kernels, g++ for the C++ host.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 4, force: bool = True) -> None:
    args = ["make", "-C", HERE, f"-j{jobs}"]
    if force:
        args.insert(1, "-B")
    subprocess.check_call(args)


if __name__ == "__main__":
    build()
