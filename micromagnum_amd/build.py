"""In-tree build of libmmhip.so (gfx950) with the Makefile in csrc/."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs=8, force=True):
    """Compile every HIP source for gfx950 and link libmmhip.so next to this file."""
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True)
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", CSRC], check=True)
    return os.path.join(_HERE, "libmmhip.so")


if __name__ == "__main__":
    print(build())
