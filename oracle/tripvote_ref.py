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


def _lib_bi():
    L = ref_cpu.lib()
    if not hasattr(L, "_bandindex_set"):
        P, I, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
        L.ora_assign_bands.argtypes = [P, I, P, P, P, D, I, P, P, P, I, I, P, P, P, P, P, P, P, P]
        L.ora_assign_bands.restype = None
        L.ora_quest_weights.argtypes = [P, P, P, P, P, P, I, I, I, I, ctypes.c_int, P]
        L.ora_quest_weights.restype = None
        L.ora_quest.argtypes = [P, P, P, I, I, P, P, P, P]
        L.ora_quest.restype = None
        L.ora_calc_cm.argtypes = [P, P, P, P, I, I, I, P]
        L.ora_calc_cm.restype = None
        L._bandindex_set = True
    return L


def assign_bands(completelib, angTol, n_band_early, bandnorms, rank_arg, bandFam):
    """_assign_bands_nb2 (tripletvote.py:1483-1688): (fit, polematch, polevalid, nMatch,
    whGood, ij, R, fitb)."""
    L = _lib_bi()
    pflt = np.ascontiguousarray(completelib["polesCart"], dtype=np.float32)
    tab = np.ascontiguousarray(completelib["angTable"], dtype=np.float64)
    fi = np.ascontiguousarray(completelib["famIndex"], dtype=np.int64)
    nf = np.ascontiguousarray(completelib["nFamily"], dtype=np.int64)
    bn = np.ascontiguousarray(bandnorms, dtype=np.float32)
    npats, nb = bn.shape[:2]
    ra = np.ascontiguousarray(rank_arg, dtype=np.int64)
    bf = np.ascontiguousarray(bandFam, dtype=np.int32)
    fit = np.zeros(npats, np.float32)
    pm = np.zeros((npats, nb), np.int64)
    pv = np.zeros((npats, nb), np.uint8)
    nm = np.zeros(npats, np.int64)
    wg = np.zeros((npats, nb), np.int64)
    ij = np.zeros((npats, 4), np.int64)
    R = np.zeros((npats, 3, 3), np.float32)
    fitb = np.zeros((npats, nb), np.float32)
    L.ora_assign_bands(_p(pflt), pflt.shape[0], _p(tab), _p(fi), _p(nf), float(angTol), int(n_band_early), _p(bn),
                       _p(ra), _p(bf), npats, nb, _p(fit), _p(pm), _p(pv), _p(nm), _p(wg), _p(ij), _p(R), _p(fitb))
    return fit, pm, pv, nm, wg, ij, R, fitb


def quest_weights(libFamID, acc, acc_nw, polematch, polevalid, band_intensity, nfit=6, simpleqweights=False):
    """_calc_quest_weights (tripletvote.py:855-897): (npats, nbands) float32."""
    L = _lib_bi()
    acc = np.ascontiguousarray(acc, dtype=np.float32)
    npats, nfam, nb = acc.shape
    w = np.zeros((npats, nb), np.float32)
    L.ora_quest_weights(_p(np.ascontiguousarray(libFamID, dtype=np.int32)), _p(acc),
                        _p(np.ascontiguousarray(acc_nw, dtype=np.float32)),
                        _p(np.ascontiguousarray(polematch, dtype=np.int64)),
                        _p(np.ascontiguousarray(polevalid, dtype=np.uint8)),
                        _p(np.ascontiguousarray(band_intensity, dtype=np.float32)), npats, nfam, nb, int(nfit),
                        int(bool(simpleqweights)), _p(w))
    return w


def quest(polescart, bandnorms, weights):
    """_orientation_quest_nb (tripletvote.py:920-1018) on matched poles (npats, nb, 3),
    band normals (npats, nb, 3) and normalised weights: (quat, lambda, mean dot, dots)."""
    L = _lib_bi()
    pc = np.ascontiguousarray(polescart, dtype=np.float64)
    bn = np.ascontiguousarray(bandnorms, dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    npats, nb = bn.shape[:2]
    q = np.zeros((npats, 4)); lam = np.zeros(npats); mad = np.zeros(npats); bd = np.zeros((npats, nb))
    L.ora_quest(_p(pc), _p(bn), _p(w), npats, nb, _p(q), _p(lam), _p(mad), _p(bd))
    return q, lam, mad, bd


def calc_cm(acc, polematch, polevalid, libFamID):
    """_calc_cm (tripletvote.py:2056-2075): (npats,) float32."""
    L = _lib_bi()
    acc = np.ascontiguousarray(acc, dtype=np.float32)
    npats, nfam, nb = acc.shape
    cm = np.zeros(npats, np.float32)
    L.ora_calc_cm(_p(acc), _p(np.ascontiguousarray(polematch, dtype=np.int64)),
                  _p(np.ascontiguousarray(polevalid, dtype=np.uint8)),
                  _p(np.ascontiguousarray(libFamID, dtype=np.int32)), npats, nfam, nb, _p(cm))
    return cm


def bandindex(indexer, band_norms, band_intensity=None):
    """BandIndexer.bandindex (tripletvote.py:492-633) with the vote, assignment, QUEST and
    confidence metric restated: (avequat, fit, cm2, polematch, nMatch, ij, acc_correct,
    bandfit) as the reference returns them.  ``indexer``: a pyebsdindex_amd BandIndexer
    (its library and settings)."""
    bn = np.asarray(band_norms, dtype=np.float32)
    bn = bn[None] if bn.ndim == 2 else bn
    npats, nb = bn.shape[:2]
    inten = np.ones((npats, nb)) if band_intensity is None else np.atleast_2d(band_intensity)
    accW, fam, rank, cm, acc = tripvote(bn, inten, indexer.lut, indexer.angTol, indexer.angtriplets["angles"],
                                        indexer.angtriplets["familyid"], indexer.npolefamilies)
    rank_arg = np.argsort(rank, axis=1, kind="stable")
    lib = indexer.completelib
    fit, pm, pv, nMatch, whGood, ij, R, fitb = assign_bands(lib, indexer.angTol, indexer.nband_earlyexit, bn,
                                                            rank_arg, fam)
    famid = lib["familyid"]
    acc_correct = np.sum(((pv > 0) & (famid[pm.clip(0)] == fam)).astype(int), axis=1).astype(np.int32)
    if indexer.high_fidelity:
        w = quest_weights(famid, accW, acc, pm, pv, inten, nfit=6, simpleqweights=indexer.simpleqweights)
        wn = w.astype(np.float64)
        wn /= np.maximum(np.sum(wn, axis=1), 1e-12).reshape(-1, 1)
        pc = np.asarray(lib["polesCart"][pm.clip(0)], dtype=np.float64)
        q, lam, mad, bd = quest(pc, bn.astype(np.float64), wn)
        bandfit = np.arccos(np.clip(bd, -1.0, 1.0)) * (180.0 / np.pi)
        fit = np.arccos(np.clip(mad, -1.0, 1.0)) * (180.0 / np.pi)
    else:
        raise NotImplementedError("high_fidelity=False (rotation from the TRIAD matrix) is not restated")
    cm2 = calc_cm(accW, pm, pv, famid)
    return q, fit, cm2, pm, nMatch, ij, acc_correct, bandfit
