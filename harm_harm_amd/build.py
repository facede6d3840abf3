"""Build step for libharmstep.so (hipcc for gfx950 + gcc for the C host)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build_library(force=True, jobs=8):
    cmd = ["make", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.run(cmd, check=True)
    return os.path.join(_HERE, "libharmstep.so")
