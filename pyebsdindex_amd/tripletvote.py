"""Triplet voting of band normals against a phase's angle-triplet library on the device.

PyEBSDIndex indexes a pattern by voting: every triplet of detected bands gives three
inter-band angles, each library triplet of pole-family angles within ``angTol`` of them
votes for a family on each of the three bands, and the votes rank the bands and name
their most likely family (``BandIndexer._tripvote_numba``, pyebsdindex/tripletvote.py:
1019-1207, called from ``BandIndexer.bandindex`` at :523-524).  This module holds

* ``tripvote_device``: the drop-in for ``_tripvote_numba`` -- same arguments, same five
  outputs -- computed by ``libebsd_tripvote.so`` (one wavefront per pattern; the library
  rows are tested 64 at a time and the matches applied in library order, so the float
  sums follow the reference's order);
* ``triplet_library`` / ``addphase``: the phase library (pole families, pair angles,
  angle triplets) of a cubic phase of Laue class m-3m, the input the vote needs
  (tripletvote.py:320-491, :70-140 for the FCC/BCC defaults);
* ``BandIndexer``: holds a phase's library on the host and the device and votes.

Band assignment, the QUEST orientation fit and the confidence metrics that follow the
vote in ``bandindex`` (tripletvote.py:548-633) are not part of this module.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from ._hip import check_tripvote, load_tripvote

RADEG = 180.0 / np.pi

FCC_POLES = np.array([[0, 0, 2], [1, 1, 1], [0, 2, 2], [1, 1, 3]], dtype=np.int32)
BCC_POLES = np.array([[0, 1, 1], [0, 0, 2], [1, 1, 2], [0, 1, 3]], dtype=np.int32)


def triplet_lut() -> np.ndarray:
    """The (3, 3, 3, 3) table the vote uses to carry family ids through a sort of a
    triplet's angles (tripletvote.py:207-213).  A triplet lists its angles as edges
    (AB, AC, BC) of the pole triangle and its family ids as vertices (A, B, C); the vertex
    at place v sits opposite edge 2 - v.  When the edges are reordered as srt (new edge m
    = old edge srt[m]), the vertex at new place v is the old vertex 2 - srt[2 - v]; the
    entry lut[:, s0, s1, s2] is that vertex order for each permutation (s0, s1, s2)."""
    lut = np.zeros((3, 3, 3, 3), dtype=np.int64)
    for s0 in range(3):
        for s1 in range(3):
            s2 = 3 - s0 - s1
            if s0 == s1 or not 0 <= s2 < 3 or s2 in (s0, s1):
                continue
            s = (s0, s1, s2)
            lut[:, s0, s1, s2] = [2 - s[2 - v] for v in range(3)]
    return lut


def _cubic_rotations() -> np.ndarray:
    """The 24 proper rotations of the cube as integer matrices (signed permutations with
    determinant +1), the proper operations of Laue class m-3m."""
    out = []
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        for signs in np.ndindex(2, 2, 2):
            m = np.zeros((3, 3), dtype=np.int64)
            for r in range(3):
                m[r, perm[r]] = -1 if signs[r] else 1
            if round(np.linalg.det(m)) == 1:
                out.append(m)
    return np.array(out)


def _unique_hkl(poles: np.ndarray) -> np.ndarray:
    """Distinct integer poles, ordered by the key h + b k + b^2 l of the shifted indices
    (tripletvote.py:645-687 with reduceInversion=False), then reversed as build_trip_lib
    does (:350-351)."""
    p = np.asarray(poles, dtype=np.int64).reshape(-1, 3)
    q = p - p.min()
    b = q.max() + 1
    key = q @ np.array([1, b, b * b], dtype=np.int64)
    _, first = np.unique(key, return_index=True)
    return p[first][::-1]


def _pole_cos(p1: np.ndarray, p2: np.ndarray) -> np.ndarray:
    """Cosine of the angle between integer cubic poles, float64 (tripletvote.py:689-707
    with the identity reciprocal metric of a cubic cell of unit edge)."""
    a = np.asarray(p1, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(p2, dtype=np.float64).reshape(-1, 3)
    dot = a @ b.T
    na = np.sqrt(np.einsum("ij,ij->i", a, a)).reshape(-1, 1)
    nb = np.sqrt(np.einsum("ij,ij->i", b, b)).reshape(1, -1)
    return np.clip(dot / (na @ nb), -1.0, 1.0)


def _normalise_families(polefamilies) -> np.ndarray:
    """Drop zero poles and poles parallel or antiparallel to an earlier one
    (tripletvote.py:245-266)."""
    poles = np.atleast_2d(np.asarray(polefamilies, dtype=np.float64))
    poles = poles[np.abs(poles).max(axis=1) > 1e-6]
    if poles.shape[0] == 0:
        raise ValueError("no non-zero pole families")
    unit = poles / np.sqrt((poles ** 2).sum(-1))[:, None]
    keep = np.ones(poles.shape[0], dtype=bool)
    dots = np.abs(unit @ unit.T)
    for i in range(poles.shape[0]):
        keep[i + 1:][dots[i, i + 1:] > 0.99999] = False
    return np.rint(poles[keep] * (1.0 + 1e-6)).astype(np.int64)


def triplet_library(polefamilies) -> dict:
    """Pair angles and angle triplets of a cubic (m-3m) phase (tripletvote.py:320-491).

    Returns a dict with
      ``polefamilies`` (nfam, 3) int, ``nfam``;
      ``families`` list of (n_i, 3) int arrays, every symmetric variant of each family;
      ``angpairs``: ``angles`` (npair,) float64 degrees and ``familyid`` (npair, 2) int --
        for families i <= j, each distinct angle (to 0.01 degree) above 1 degree between
        family i's listed pole and the variants of family j;
      ``angtriplets``: ``angles`` (ntrip, 3) float64, ascending in each row, and
        ``familyid`` (ntrip, 3) int64, the family at each vertex, in the reference's row
        order (duplicates removed, rows ordered by their family-id key)."""
    poles = _normalise_families(polefamilies)
    nfam = poles.shape[0]
    rots = _cubic_rotations()
    families = [_unique_hkl(np.einsum("rij,j->ri", rots, p)) for p in poles]

    angs, fid = [], []
    for i in range(nfam):
        for j in range(i, nfam):
            c = np.abs(_pole_cos(poles[i], families[j]).ravel())
            cents = np.round(np.arccos(c) * RADEG * 100).astype(np.int32)
            u = np.unique(cents) / 100.0
            for a in u[u > 1.0]:
                angs.append(a)
                fid.append((i, j))
    angs = np.array(angs, dtype=np.float64)
    fid = np.array(fid, dtype=np.int64).reshape(-1, 2)

    # every (A-B, A-C, B-C) combination: A in family i, (B, C) from i's pair list with
    # j <= k, B-C any listed angle between the families of B and C (tripletvote.py:418-447)
    rows_a, rows_i = [], []
    for i in range(nfam):
        own = np.nonzero(fid[:, 0] == i)[0]
        for x, jx in enumerate(own):
            for kx in own[x:]:
                f1, f2 = fid[jx, 1], fid[kx, 1]
                for q in np.nonzero((fid[:, 0] == f1) & (fid[:, 1] == f2))[0]:
                    rows_a.append((angs[jx], angs[kx], angs[q]))
                    rows_i.append((i, f1, f2))
    lib_a = np.array(rows_a, dtype=np.float64).reshape(-1, 3)
    lib_i = np.array(rows_i, dtype=np.int64).reshape(-1, 3)

    # ascending angles in each row, the family ids carried along (tripletvote.py:713-735)
    lut = triplet_lut()
    for r in range(lib_a.shape[0]):
        s = np.argsort(lib_a[r])
        lib_a[r] = lib_a[r][s]
        lib_i[r] = lib_i[r][lut[:, s[0], s[1], s[2]]]
    # duplicates (same rounded angle sum and ids) removed, then rows ordered by the id key
    # (tripletvote.py:737-755); the same numpy sorts, so the same row order
    ang_key = np.sum(np.round(lib_a * 100), axis=1).astype(np.longlong)
    base = np.longlong(lib_i.max() + 1)
    id_key = lib_i.dot(np.array([1, base, base ** 2]))
    digits = np.ceil(np.log10(id_key.max()))
    digits = np.where(digits > 2, digits, 2)
    uid = ang_key * 10 ** digits + id_key
    _, first = np.unique(uid, return_index=True)
    lib_a, lib_i, id_key = lib_a[first], lib_i[first], id_key[first]
    order = np.argsort(id_key)
    return {"polefamilies": poles, "nfam": nfam, "families": families,
            "angpairs": {"angles": angs, "familyid": fid},
            "angtriplets": {"angles": lib_a[order], "familyid": lib_i[order]}}


def row_index(tripAngles, angTol: float, bin_width: float = 0.5, max_angle: float = 90.0):
    """Index of library rows by their first (smallest) angle, for the device vote: bin b
    covers smallest triplet angles x0 in [b w, (b+1) w) (the first bin down to -inf, the
    last up to +inf) and lists, ascending, every row whose first angle is within angTol
    (plus 1e-3 degree of slack: the kernel finds the bin with a float32 product) of that
    range.  Returns (start (nbins+1,) int32, rows int32, bin_width, angTol): the index
    only serves votes whose tolerance is at most the one it was built for."""
    a0 = np.asarray(tripAngles, dtype=np.float64).reshape(-1, 3)[:, 0]
    nbins = int(math.ceil(max_angle / bin_width)) + 1
    lo = np.arange(nbins) * bin_width - angTol - 1e-3
    hi = (np.arange(nbins) + 1) * bin_width + angTol + 1e-3
    lo[0], hi[-1] = -np.inf, np.inf
    lists = [np.nonzero((a0 >= lo[b]) & (a0 <= hi[b]))[0] for b in range(nbins)]
    start = np.zeros(nbins + 1, dtype=np.int32)
    start[1:] = np.cumsum([len(x) for x in lists])
    rows = np.concatenate(lists).astype(np.int32) if lists else np.zeros(0, np.int32)
    return start, rows, float(bin_width), float(angTol)


def addphase(libtype: str = "FCC", angTol: float = 2.0, polefamilies=None):
    """A BandIndexer for an FCC (space group 225) or BCC (229) phase with the reference's
    default pole families (tripletvote.py:70-140), or cubic families of one's own."""
    t = str(libtype).upper()
    if polefamilies is None:
        if t == "FCC":
            polefamilies = FCC_POLES
        elif t == "BCC":
            polefamilies = BCC_POLES
        else:
            raise NotImplementedError(f"libtype {libtype!r}: only cubic (m-3m) phases, FCC or BCC, are built")
    return BandIndexer(phasename=t, polefamilies=polefamilies, angTol=angTol)


class BandIndexer:
    """A cubic phase's triplet library and its vote on the device (the library and voting
    half of tripletvote.BandIndexer, tripletvote.py:171-491, :1019-1207)."""

    def __init__(self, phasename: str = "", polefamilies=FCC_POLES, angTol: float = 2.0,
                 nband_earlyexit: int = 8):
        self.phasename = str(phasename)
        self.angTol = float(angTol)
        self.nband_earlyexit = int(nband_earlyexit)
        self.lut = triplet_lut()
        lib = triplet_library(polefamilies)
        self.polefamilies = lib["polefamilies"]
        self.npolefamilies = lib["nfam"]
        self.angpairs = lib["angpairs"]
        self.angtriplets = lib["angtriplets"]
        self.families = lib["families"]
        self._dev = {}

    def _device_library(self, device):
        import torch
        key = (str(device), self.angTol)
        if key not in self._dev:
            a = torch.as_tensor(np.ascontiguousarray(self.angtriplets["angles"], dtype=np.float64), device=device)
            f = torch.as_tensor(np.ascontiguousarray(self.angtriplets["familyid"], dtype=np.int32), device=device)
            start, rows, bw, tol = row_index(self.angtriplets["angles"], self.angTol)
            idx = (torch.as_tensor(start, device=device), torch.as_tensor(rows, device=device), bw, tol) \
                if rows.size else None
            self._dev[key] = (a, f, idx)
        return self._dev[key]

    def tripvote(self, bandnorms, band_intensity=None, stream=None, use_index=True):
        """Votes of (npats, nbands, 3) float32 band normals (device tensor); returns the
        five outputs of _tripvote_numba as device tensors.  The library and its row index
        stay on the device between calls."""
        a, f, idx = self._device_library(bandnorms.device)
        return tripvote_device(bandnorms, band_intensity, self.lut, self.angTol, a, f,
                               self.npolefamilies, stream=stream, index=idx if use_index else None)

    def bandindex(self, *args, **kwargs):
        raise NotImplementedError("band assignment and the QUEST orientation fit that follow the vote "
                                  "(tripletvote.py:548-633) are not built on the device")


def tripvote_device(bandnorms, band_intensity, LUT, angTol, tripAngles, tripID, nfam, stream=None,
                    index="auto"):
    """Device counterpart of BandIndexer._tripvote_numba (tripletvote.py:1019-1207).

    bandnorms: (npats, nbands, 3) or (nbands, 3) float32 unit band normals on the device.
    band_intensity: (npats, nbands) or None (all ones); a band below 1e-6 takes no part.
    LUT: the (3, 3, 3, 3) triplet table (must equal triplet_lut()).
    angTol: degrees.  tripAngles (ntrip, 3) float64 ascending per row, tripID (ntrip, 3)
    family ids in [0, nfam): the phase library, numpy or device tensors.
    Returns (accumulatorW, bandFam, bandRank, band_cm, accumulator) as device tensors of
    shapes (npats, nfam, nbands) float32, (npats, nbands) int32, (npats, nbands) float32
    x2, (npats, nfam, nbands) float32 -- the reference's output order (weighted votes
    first, vote counts last).
    index: "auto" builds the row index of row_index() when tripAngles is a host array
    (the outputs do not depend on it), None tests every row, or the (start, rows, width,
    built_tol) of row_index() with start and rows as device tensors; an index built for a
    tolerance below angTol would drop matching rows and is refused."""
    import torch
    lib = load_tripvote()
    if not isinstance(bandnorms, torch.Tensor) or bandnorms.device.type != "cuda":
        raise TypeError("bandnorms must be a device tensor")
    if bandnorms.dtype != torch.float32:
        raise TypeError("bandnorms must be float32 (radon2pole's output type)")
    bn = bandnorms if bandnorms.dim() == 3 else bandnorms.unsqueeze(0)
    if bn.dim() != 3 or bn.shape[2] != 3:
        raise ValueError(f"bandnorms must be (npats, nbands, 3), got {tuple(bandnorms.shape)}")
    bn = bn.contiguous()
    dev = bn.device
    npats, nb = int(bn.shape[0]), int(bn.shape[1])
    if not np.array_equal(np.asarray(LUT), triplet_lut()):
        raise ValueError("LUT differs from the triplet table of BandIndexer (tripletvote.py:207-213)")
    if band_intensity is None:
        inten = torch.ones((npats, nb), dtype=torch.float32, device=dev)
    else:
        inten = torch.as_tensor(band_intensity, device=dev)
        if inten.dim() == 1:
            inten = inten.unsqueeze(0)
        inten = inten.to(torch.float32).contiguous()
        if tuple(inten.shape) != (npats, nb):
            raise ValueError(f"band_intensity must be ({npats}, {nb}), got {tuple(inten.shape)}")
    if isinstance(tripID, np.ndarray) and tripID.size and (tripID.min() < 0 or tripID.max() >= int(nfam)):
        raise ValueError(f"tripID holds families outside [0, {int(nfam)})")
    if isinstance(index, str):
        if index != "auto":
            raise ValueError("index must be 'auto', None or the 4-tuple of row_index()")
        index = None
        if isinstance(tripAngles, np.ndarray):
            st, rw, bw, tol = row_index(tripAngles, float(angTol))
            if rw.size:
                index = (torch.as_tensor(st, device=dev), torch.as_tensor(rw, device=dev), bw, tol)
    elif index is not None:
        if len(index) != 4:
            raise ValueError("index must be the (start, rows, width, built_tol) 4-tuple of row_index()")
        if float(index[3]) < float(angTol):
            raise ValueError(f"row index built for angTol {float(index[3])} cannot serve angTol {float(angTol)}")
    ta = torch.as_tensor(tripAngles, device=dev).to(torch.float64).contiguous()
    ti = torch.as_tensor(tripID, device=dev).to(torch.int32).contiguous()
    if ta.dim() != 2 or ta.shape[1] != 3 or tuple(ti.shape) != tuple(ta.shape):
        raise ValueError("tripAngles and tripID must both be (ntrip, 3)")
    nfam = int(nfam)
    accW = torch.empty((npats, nfam, nb), dtype=torch.float32, device=dev)
    acc = torch.empty_like(accW)
    fam = torch.empty((npats, nb), dtype=torch.int32, device=dev)
    rank = torch.empty((npats, nb), dtype=torch.float32, device=dev)
    cm = torch.empty_like(rank)
    from ._hip import keep_for_stream, launch_stream
    stream = launch_stream(stream, dev)
    P = ctypes.c_void_p
    rc = lib.ebsd_tripvote(P(bn.data_ptr()), P(inten.data_ptr()), ctypes.c_int64(npats), ctypes.c_int(nb),
                           P(ta.data_ptr()), P(ti.data_ptr()), ctypes.c_int(int(ta.shape[0])),
                           ctypes.c_int(nfam), ctypes.c_double(float(angTol)),
                           P(index[0].data_ptr() if index is not None else None),
                           P(index[1].data_ptr() if index is not None else None),
                           ctypes.c_int(int(index[0].numel()) - 1 if index is not None else 0),
                           ctypes.c_double(index[2] if index is not None else 0.0),
                           P(accW.data_ptr()), P(fam.data_ptr()), P(rank.data_ptr()), P(cm.data_ptr()),
                           P(acc.data_ptr()), P(stream.cuda_stream))
    check_tripvote(lib, rc)
    keep_for_stream(stream, bn, inten, ta, ti, accW, fam, rank, cm, acc,
                    *(index[:2] if index is not None else ()))
    return accW, fam, rank, cm, acc


def bandangles(bandnorms: np.ndarray) -> np.ndarray:
    """Host helper: the (npats, nbands, nbands) inter-band angles in degrees."""
    bn = np.asarray(bandnorms, dtype=np.float64)
    c = np.clip(np.abs(np.einsum("pik,pjk->pij", bn, bn)), -1.0, 1.0)
    return np.degrees(np.arccos(c))


__all__ = ["triplet_lut", "triplet_library", "addphase", "BandIndexer", "tripvote_device", "bandangles",
           "FCC_POLES", "BCC_POLES"]
