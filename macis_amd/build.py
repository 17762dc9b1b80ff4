"""Build recipe: compiles every HIP translation unit under macis_amd/csrc
for gfx950 into macis_amd/libmacis_hip.so (in-tree, so it travels with the
repository), plus the C++ API tests under macis_amd/cpp_tests when present."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = True) -> str:
    args = ["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    cpp = os.path.join(HERE, "cpp_tests")
    if os.path.exists(os.path.join(cpp, "Makefile")):
        args = ["make", "-C", cpp, f"-j{jobs}"] + (["-B"] if force else [])
        subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return os.path.join(HERE, "libmacis_hip.so")


if __name__ == "__main__":
    print(build())
