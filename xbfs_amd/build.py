"""Builds libxbfs.so and ./bfs in-tree (hipcc for gfx950, g++ for host code)."""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))


def build(force=True, jobs=4):
    args = ["make", "-C", os.path.join(PKG, "csrc"), f"-j{jobs}"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(PKG, "libxbfs.so")


if __name__ == "__main__":
    build()
