"""Build the CPU oracle (test infrastructure only) into oracle/liboracle.so.

The reference itself cannot be compiled here: every translation unit on its
hot path includes HighFive (src/hdf_interface.hpp:12), a header library this
image does not have, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-Wall", "-Wextra", "-o", LIB, os.path.join(HERE, "oracle.c"), "-lm"]
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + p.stdout)
    return LIB


if __name__ == "__main__":
    print(build())
    sys.exit(0)
