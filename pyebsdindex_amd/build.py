"""Build the gfx950 HIP libraries in-tree.

``python -m pyebsdindex_amd.build`` (or ``__graft_entry__.build()``) runs hipcc on
``csrc/ebsd_radon.hip`` (band detection -> ``libebsd_radon.so``) and
``csrc/ebsd_nlpar.hip`` (NLPAR -> ``libebsd_nlpar.so``) and ``csrc/ebsd_tripvote.hip``
(triplet voting -> ``libebsd_tripvote.so``), plus the measurement-only
``csrc/ebsd_ldsprobe.hip`` (the LDS read and VALU issue ceilings -> ``libebsd_ldsprobe.so``, used by bench.py).  All are plain C-ABI shared objects (no torch
symbols); ``pyebsdindex_amd._hip`` loads them with ctypes.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "ebsd_radon.hip")
LIB = os.path.join(HERE, "libebsd_radon.so")
NLPAR_SRC = os.path.join(HERE, "csrc", "ebsd_nlpar.hip")
NLPAR_LIB = os.path.join(HERE, "libebsd_nlpar.so")
TRIPVOTE_SRC = os.path.join(HERE, "csrc", "ebsd_tripvote.hip")
TRIPVOTE_LIB = os.path.join(HERE, "libebsd_tripvote.so")
PROBE_SRC = os.path.join(HERE, "csrc", "ebsd_ldsprobe.hip")
PROBE_LIB = os.path.join(HERE, "libebsd_ldsprobe.so")
INCLUDE = os.path.join(ROOT, "include")

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # float contraction only where the source asks for fmaf(): keeps the f64 parity
    # arithmetic (sub-pixel fit, np.interp restatement) free of fused operations
    "-ffp-contract=off",
    "-fno-strict-aliasing",
    # no packed-FP32 VALU instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc(src: str, lib: str, verbose: bool) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = lib + ".tmp"
    cmd = [hipcc] + HIPCC_FLAGS + ["-I", INCLUDE, "-o", tmp, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


def build(verbose: bool = False) -> str:
    """Compile the libraries; returns the band-detection library's path."""
    _hipcc(PROBE_SRC, PROBE_LIB, verbose)
    _hipcc(TRIPVOTE_SRC, TRIPVOTE_LIB, verbose)
    _hipcc(NLPAR_SRC, NLPAR_LIB, verbose)
    return _hipcc(SRC, LIB, verbose)


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
