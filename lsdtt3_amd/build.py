"""Build liblsdtt3_gpu.so in-tree: ``python -m lsdtt3_amd.build``."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = False) -> str:
    if force:
        subprocess.run(["make", "-s", "-C", HERE, "clean"], check=True)
    subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}"], check=True)
    return os.path.join(HERE, "lib", "liblsdtt3_gpu.so")


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
