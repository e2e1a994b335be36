"""Build recipe for the oracle's C restatement (test infrastructure only).

``python oracle/build.py`` compiles ``oracle/csrc/ref_loops.c`` with gcc/OpenMP into
``oracle/libref_loops.so``.  The reference itself is Python + Numba (no C/C++ sources),
so there is no ``oracle/_ref`` build: the reference is read, not run.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "ref_loops.c")
LIB = os.path.join(HERE, "libref_loops.so")


def build(force: bool = True) -> str:
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-fPIC",
           "-shared", "-o", tmp, SRC, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
    sys.exit(0)
