"""Build the in-tree gfx950 library (cardiomechanics_amd/libcmg.so) with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force: bool = True, jobs: int = 8) -> str:
    cmd = ["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"]
    if force:
        cmd.append("-B")
    subprocess.check_call(cmd)
    return os.path.join(HERE, "libcmg.so")


if __name__ == "__main__":
    print(build(force="--no-force" not in sys.argv))
