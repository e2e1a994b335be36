"""Build recipe for the oracle's C restatement (test infrastructure only).

``python oracle/build.py`` compiles ``oracle/csrc/ref_loops.c`` (band detection) and
``oracle/csrc/ref_nlpar.c`` (NLPAR), ``oracle/csrc/ref_tripvote.c`` (triplet vote) and
``oracle/csrc/ref_bandindex.c`` (band assignment, QUEST, confidence metric) with gcc/OpenMP
into ``oracle/libref_loops.so``.  The reference itself is Python + Numba (no C/C++ sources),
so there is no ``oracle/_ref`` build: the reference is read, not run.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", "ref_loops.c"), os.path.join(HERE, "csrc", "ref_nlpar.c"),
        os.path.join(HERE, "csrc", "ref_tripvote.c"), os.path.join(HERE, "csrc", "ref_bandindex.c")]
LIB = os.path.join(HERE, "libref_loops.so")


def build(force: bool = True) -> str:
    if not force and os.path.exists(LIB) and all(os.path.getmtime(LIB) >= os.path.getmtime(s) for s in SRCS):
        return LIB
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-fPIC",
           "-shared", "-o", tmp] + SRCS + ["-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
    sys.exit(0)
