"""Build libfcw.so in-tree (hipcc for gfx950, g++ for the host analysis)."""

import os
import subprocess
import sys

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs: int = 8, force: bool = False) -> str:
    """Compile every source of libfcw.so; returns the library path."""
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)
    return os.path.join(_HERE, "libfcw.so")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
