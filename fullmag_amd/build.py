"""Builds libfullmag_fdm.so.0 for gfx950 in-tree (fullmag_amd/lib/)."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force: bool = True, jobs: int = 8) -> str:
    args = ["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "lib", "libfullmag_fdm.so.0")
