"""Builds libagdeblend.so in-tree (hipcc for gfx950 + gcc for the C host)."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")


def build(force=True, jobs=4):
    cmd = ["make", "-C", CSRC, "-j%d" % jobs]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    lib = os.path.join(PKG, "lib", "libagdeblend.so")
    if not os.path.exists(lib):
        raise RuntimeError("build did not produce " + lib)
    return lib


if __name__ == "__main__":
    print(build(force="--no-force" not in sys.argv))
