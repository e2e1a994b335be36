"""Triplet-vote oracle -- TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``tripvote``: BandIndexer._tripvote_numba (pyebsdindex/tripletvote.py:1019-1207) restated
in C (``csrc/ref_tripvote.c``), float32 where the reference is float32; also the timed CPU
baseline of the vote.

How it is pinned: tests/test_oracle_tripvote.py -- a hand-solved three-band triangle
(which band takes which family, weights angTol for exact angles), exact cubic poles
(every band's true family has a vote from each triplet it is in), an invalid last band
(the other bands' votes equal those of the shorter band list, bit for bit), band
permutations (vote counts permute), and the library's golden pair-angle counts and
closed-form angles.

Parity unpinned: the form of the float32 inter-band cosine.  The reference takes it from a
2-D ``.dot`` (tripletvote.py:1192) that Numba hands to BLAS, whose summation order and use
of fused multiply-adds are not known here; the oracle (and the device) evaluate
((x0 y0) + (x1 y1)) + (x2 y2) without contraction.  A test places a library angle at the
tolerance edge to show that the two forms can vote differently there.
"""
import ctypes

import numpy as np

from . import ref_cpu


def _lib():
    L = ref_cpu.lib()
    if not hasattr(L, "_tripvote_set"):
        P, I = ctypes.c_void_p, ctypes.c_int64
        L.ora_tripvote.argtypes = [P, P, I, I, P, ctypes.c_double, P, P, I, I, P, P, P, P, P]
        L.ora_tripvote.restype = None
        L._tripvote_set = True
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def tripvote(bandnorms, band_intensity, LUT, angTol, tripAngles, tripID, nfam):
    """Same arguments and outputs as _tripvote_numba: (accumulatorW, bandFam, bandRank,
    band_cm, accumulator)."""
    bn = np.ascontiguousarray(bandnorms, dtype=np.float32)
    if bn.ndim == 2:
        bn = bn[None]
    npats, nb = bn.shape[:2]
    if band_intensity is None:
        inten = np.ones((npats, nb), dtype=np.float32)
    else:
        inten = np.ascontiguousarray(np.atleast_2d(band_intensity), dtype=np.float32)
    lut = np.ascontiguousarray(LUT, dtype=np.int64)
    ta = np.ascontiguousarray(tripAngles, dtype=np.float64)
    ti = np.ascontiguousarray(tripID, dtype=np.int64)
    nfam = int(nfam)
    accW = np.zeros((npats, nfam, nb), np.float32)
    acc = np.zeros_like(accW)
    fam = np.zeros((npats, nb), np.int32)
    rank = np.zeros((npats, nb), np.float32)
    cm = np.zeros_like(rank)
    _lib().ora_tripvote(_p(bn), _p(inten), npats, nb, _p(lut), float(angTol), _p(ta), _p(ti),
                        ta.shape[0], nfam, _p(accW), _p(fam), _p(rank), _p(cm), _p(acc))
    return accW, fam, rank, cm, acc
