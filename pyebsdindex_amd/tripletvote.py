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
  angle triplets, the complete pole list) of a phase of any Laue class, the input the vote
  and the band assignment need
  (tripletvote.py:320-491, :70-140 for the FCC/BCC defaults);
* ``BandIndexer``: holds a phase's library on the host and the device and votes.

``BandIndexer.bandindex`` runs the rest on the device too (``ebsd_bandindex``): band
assignment, the QUEST orientation fit and the confidence metric (tripletvote.py:526-633).
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


# --------------------------------------------------------------------------------------
# crystal geometry (pyebsdindex/crystallometry.py:63-165, crystal_sym.py:35-482): the
# Cartesian frame has a along x and c* along z; pole families are the orbits of integer
# plane normals under the proper rotations of the phase's Laue class
# --------------------------------------------------------------------------------------
HEX_POLES = np.array([[1, 0, -1, 0], [0, 0, 0, 2], [1, 0, -1, 1], [1, 0, -1, 2], [1, 1, -2, 0],
                      [1, 0, -1, 3], [1, 1, -2, 2], [2, 0, -2, 1]], dtype=np.int32)

# space group -> Laue class, as the first space group of each class (EDAX numbering of the
# classes: 1 = -1, 2 = 2/m, 22 = mmm, 4 = 4/m, 42 = 4/mmm, 3 = -3, 32 = -3m, 6 = 6/m,
# 62 = 6/mmm, 23 = m-3, 43 = m-3m; crystal_sym.py:379-406)
_LAUE_FIRST_SG = ((1, 1), (3, 2), (16, 22), (75, 4), (89, 42), (143, 3), (149, 32), (168, 6), (177, 62),
                  (195, 23), (207, 43))


def laue_class(spacegroup: int) -> int:
    """The Laue class (EDAX code) of a space group 1..230."""
    sg = int(spacegroup)
    if not 1 <= sg <= 230:
        raise ValueError(f"space group {spacegroup} outside 1..230")
    code = 1
    for first, c in _LAUE_FIRST_SG:
        if sg >= first:
            code = c
    return code


def _axis_rotation(axis, degrees) -> np.ndarray:
    """Rotation matrix about a Cartesian axis (float64)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    t = np.radians(degrees)
    K = np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])
    return np.eye(3) + np.sin(t) * K + (1.0 - np.cos(t)) * (K @ K)


def laue_rotations(code: int) -> np.ndarray:
    """The proper rotations of a Laue class as Cartesian matrices (n, 3, 3), with the axis
    conventions of crystal_sym.py:35-377: n-fold axes along z (c), the hexagonal two-folds
    every 30 degrees in the basal plane from x (a), the trigonal ones at 0, 120 and 240
    degrees, the tetragonal ones along x, y and the two diagonals, the monoclinic two-fold
    along y (b)."""
    z = (0.0, 0.0, 1.0)
    R = []
    if code in (43, 23):
        for perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)):
            for signs in np.ndindex(2, 2, 2):
                m = np.zeros((3, 3))
                for r in range(3):
                    m[r, perm[r]] = -1.0 if signs[r] else 1.0
                if round(np.linalg.det(m)) != 1:
                    continue
                cyclic = perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1))
                if code == 43 or cyclic:
                    R.append(m)
    elif code in (62, 6):
        R = [_axis_rotation(z, 60.0 * k) for k in range(6)]
        if code == 62:
            R += [_axis_rotation((np.cos(np.radians(30.0 * k)), np.sin(np.radians(30.0 * k)), 0.0), 180.0)
                  for k in range(6)]
    elif code in (32, 3):
        R = [_axis_rotation(z, 120.0 * k) for k in range(3)]
        if code == 32:
            R += [_axis_rotation((np.cos(np.radians(a)), np.sin(np.radians(a)), 0.0), 180.0) for a in (0, 120, 240)]
    elif code in (42, 4):
        R = [_axis_rotation(z, 90.0 * k) for k in range(4)]
        if code == 42:
            R += [_axis_rotation(ax, 180.0) for ax in ((1, 1, 0), (-1, 1, 0), (1, 0, 0), (0, 1, 0))]
    elif code == 22:
        R = [np.eye(3)] + [_axis_rotation(ax, 180.0) for ax in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
    elif code == 2:
        R = [np.eye(3), _axis_rotation((0, 1, 0), 180.0)]
    elif code == 1:
        R = [np.eye(3)]
    else:
        raise ValueError(f"unknown Laue class {code}")
    return np.array(R)


class Crystal:
    """Metric tensors and structure matrices of a unit cell (a, b, c in any length unit,
    angles in degrees): the standard formulas, with the reference's frame (a along x, c* along
    z) and its zeroing of entries below 1e-6 (crystallometry.py:63-165)."""

    def __init__(self, a, b, c, alpha, beta, gamma):
        ca, cb, cg = (np.cos(np.radians(x)) for x in (alpha, beta, gamma))
        sa, sb, sg = (np.sin(np.radians(x)) for x in (alpha, beta, gamma))
        tg = np.tan(np.radians(gamma))
        G = np.array([[a * a, a * b * cg, a * c * cb], [a * b * cg, b * b, b * c * ca], [a * c * cb, b * c * ca, c * c]])
        det = (a * b * c) ** 2 * (1.0 - ca ** 2 - cb ** 2 - cg ** 2 + 2.0 * ca * cb * cg)
        V = np.sqrt(det)
        Gs = np.array([[(b * c * sa) ** 2, a * b * c * c * (ca * cb - cg), a * b * b * c * (cg * ca - cb)],
                       [a * b * c * c * (ca * cb - cg), (a * c * sb) ** 2, a * a * b * c * (cb * cg - ca)],
                       [a * b * b * c * (cg * ca - cb), a * a * b * c * (cb * cg - ca), (a * b * sg) ** 2]]) / det
        # direct structure matrix: columns are a, b, c in Cartesian coordinates
        D = np.array([[a, b * cg, c * cb], [0.0, b * sg, -c * (cb * cg - ca) / sg], [0.0, 0.0, V / (a * b * sg)]])
        # reciprocal structure matrix: columns are a*, b*, c* (times the cell volume convention)
        Rs = np.array([[1.0 / a, 0.0, 0.0], [-1.0 / (a * tg), 1.0 / (b * sg), 0.0],
                       [b * c * (cg * ca - cb) / (V * sg), a * c * (cb * cg - ca) / (V * sg), a * b * sg / V]])
        for m in (G, Gs, D, Rs):
            m[np.abs(m) < 1e-6] = 0.0
        self.metricTensor = G
        self.reciprocalMetricTensor = Gs
        self.directStructureMatrix = D
        self.reciprocalStructureMatrix = Rs
        self.invReciprocalStructureMatrix = D.T.copy()
        self.volume = V


def hex4_to_hex3(poles) -> np.ndarray:
    """Four-index (h k i l) plane normals -> three-index (h k l) (crystal_sym.py:469-473)."""
    p = np.asarray(poles).reshape(-1, 4)
    return p[:, [0, 1, 3]]


def hex3_to_hex4(poles) -> np.ndarray:
    """(h k l) -> (h k -(h+k) l) (crystal_sym.py:475-482)."""
    p = np.asarray(poles, dtype=np.float64).reshape(-1, 3)
    return np.stack([p[:, 0], p[:, 1], -(p[:, 0] + p[:, 1]), p[:, 2]], axis=1)


def _unique_hkl(poles: np.ndarray) -> np.ndarray:
    """Distinct poles, ordered by the key h + b k + b^2 l of the rounded, shifted indices
    (tripletvote.py:645-687 with reduceInversion=False), then reversed as build_trip_lib
    does (:350-351)."""
    p = np.asarray(poles, dtype=np.float64).reshape(-1, 3)
    q = np.rint(p).astype(np.int64)
    q = q - q.min()
    b = q.max() + 1
    key = q @ np.array([1, b, b * b], dtype=np.int64)
    _, first = np.unique(key, return_index=True)
    return p[first][::-1]


def _pole_cos(p1: np.ndarray, p2: np.ndarray, rmt=None) -> np.ndarray:
    """Cosine of the angle between plane normals given by Miller indices, through the
    reciprocal metric tensor (tripletvote.py:689-707), float64, clipped to [-1, 1]."""
    G = np.eye(3) if rmt is None else np.asarray(rmt, dtype=np.float64)
    a = np.asarray(p1, dtype=np.float64).reshape(-1, 3)
    b = np.asarray(p2, dtype=np.float64).reshape(-1, 3)
    dot = a @ G @ b.T
    na = np.sqrt(np.einsum("ij,jk,ik->i", a, G, a)).reshape(-1, 1)
    nb = np.sqrt(np.einsum("ij,jk,ik->i", b, G, b)).reshape(1, -1)
    return np.clip(dot / (na @ nb), -1.0, 1.0)


def _normalise_families(polefamilies) -> np.ndarray:
    """Drop zero poles and poles parallel or antiparallel (in index space) to an earlier one
    (tripletvote.py:245-266); any index count (3 or 4)."""
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


def _family(pole, rots, crystal) -> np.ndarray:
    """The symmetric variants of one pole (Miller indices, float64): rotated in the
    Cartesian frame and mapped back, snapped to integers where they are integers to 1e-6
    (every consistent cell and Laue class), made distinct and ordered as _unique_hkl."""
    cart = crystal.reciprocalStructureMatrix @ np.asarray(pole, dtype=np.float64)
    var = np.einsum("rij,j->ri", rots, cart) @ crystal.invReciprocalStructureMatrix.T
    near = np.abs(var - np.rint(var)) < 1e-6
    var = np.where(near, np.rint(var), var)
    return _unique_hkl(var)


def triplet_library(polefamilies, spacegroup: int = 225, latticeparameter=(1.0, 1.0, 1.0, 90.0, 90.0, 90.0)) -> dict:
    """Pair angles, angle triplets and the complete pole list of a phase (tripletvote.py:
    320-491), for any Laue class.

    ``polefamilies``: (n, 3) Miller indices, or (n, 4) Miller-Bravais for the hexagonal Laue
    classes.  Returns a dict with
      ``polefamilies`` the distinct families as given (four-index for hexagonal input), ``nfam``;
      ``families`` list of (n_i, 3) arrays, every symmetric variant of each family;
      ``angpairs``: ``angles`` (npair,) float64 degrees and ``familyid`` (npair, 2) int --
        for families i <= j, each distinct angle (to 0.01 degree) above 1 degree between
        family i's listed pole and the variants of family j;
      ``angtriplets``: ``angles`` (ntrip, 3) float64, ascending in each row, and
        ``familyid`` (ntrip, 3) int64, the family at each vertex, in the reference's row
        order (duplicates removed, rows ordered by their family-id key);
      ``completelib``: ``poles`` (npoles, 3) every variant of every family, ``polesCart``
        unit Cartesian normals, ``familyid`` per pole, ``angTable`` (npoles, npoles) degrees,
        ``nFamily`` variants per family, ``famIndex`` first pole of each family."""
    code = laue_class(spacegroup)
    lp = np.asarray(latticeparameter, dtype=np.float64).reshape(6)
    crystal = Crystal(*lp)
    given = _normalise_families(polefamilies)
    hexagonal = code in (62, 6)
    poles = hex4_to_hex3(given) if (hexagonal and given.shape[-1] == 4) else given.reshape(-1, 3)
    poles = np.asarray(poles, dtype=np.int64)
    nfam = poles.shape[0]
    rots = laue_rotations(code)
    rmt = crystal.reciprocalMetricTensor
    families = [_family(p, rots, crystal) for p in poles]

    angs, fid = [], []
    for i in range(nfam):
        for j in range(i, nfam):
            c = np.abs(_pole_cos(poles[i], families[j], rmt).ravel())
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
        s_ = np.argsort(lib_a[r])
        lib_a[r] = lib_a[r][s_]
        lib_i[r] = lib_i[r][lut[:, s_[0], s_[1], s_[2]]]
    # duplicates (same rounded angle sum and ids) removed, then rows ordered by the id key
    # (tripletvote.py:737-755); the same numpy sorts, so the same row order
    if lib_a.shape[0]:
        ang_key = np.sum(np.round(lib_a * 100), axis=1).astype(np.longlong)
        base = np.longlong(lib_i.max() + 1)
        id_key = lib_i.dot(np.array([1, base, base ** 2]))
        digits = np.ceil(np.log10(max(int(id_key.max()), 1)))
        digits = np.where(digits > 2, digits, 2)
        uid = ang_key * 10 ** digits + id_key
        _, first = np.unique(uid, return_index=True)
        lib_a, lib_i, id_key = lib_a[first], lib_i[first], id_key[first]
        order = np.argsort(id_key)
        lib_a, lib_i = lib_a[order], lib_i[order]

    # the complete pole list for band assignment (tripletvote.py:463-477)
    allp = np.concatenate(families)
    nF = np.array([f.shape[0] for f in families], dtype=np.int64)
    famindex = np.concatenate(([0], np.cumsum(nF)))[:-1].astype(np.int64)
    cart = allp @ crystal.reciprocalStructureMatrix.T
    cart /= np.linalg.norm(cart, axis=1, keepdims=True)
    complete = {"poles": allp, "polesCart": cart,
                "familyid": np.repeat(np.arange(nfam, dtype=np.int32), nF),
                "angTable": np.arccos(_pole_cos(allp, allp, rmt)) * RADEG,
                "nFamily": nF, "famIndex": famindex}
    shown = hex3_to_hex4(poles) if hexagonal else poles
    return {"polefamilies": shown, "nfam": nfam, "families": families, "lauecode": code, "crystal": crystal,
            "angpairs": {"angles": angs, "familyid": fid},
            "angtriplets": {"angles": lib_a, "familyid": lib_i},
            "completelib": complete}


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


def addphase(libtype=None, phasename=None, spacegroup=None, latticeparameter=None, polefamilies=None,
             nband_earlyexit=10, pointgroup=None, pointgroupid=None, angTol: float = 2.0):
    """A BandIndexer for a phase (tripletvote.py:69-168): ``libtype`` 'FCC' (space group 225),
    'BCC' (229) or 'HCP' (194, c/a 1.63) with the reference's default pole families and an
    early exit at 8 matched bands for the cubic ones; or a space group, lattice parameters
    (a, b, c, alpha, beta, gamma) and pole families of one's own (an FCC library when no
    space group is given).  ``angTol`` (degrees) is the vote and assignment tolerance."""
    if libtype is not None:
        t = str(libtype).upper()
        if t == "FCC":
            nband_earlyexit, spacegroup = 8, spacegroup or 225
            phasename = phasename or "FCC"
            default = FCC_POLES
        elif t == "BCC":
            nband_earlyexit, spacegroup = 8, spacegroup or 229
            phasename = phasename or "BCC"
            default = BCC_POLES
        elif t == "HCP":
            spacegroup = spacegroup or 194
            phasename = phasename or "HCP"
            default = HEX_POLES
            if latticeparameter is None:
                latticeparameter = (1.0, 1.0, 1.63, 90.0, 90.0, 120.0)
        else:
            raise ValueError(f"libtype {libtype!r}: FCC, BCC or HCP")
        if latticeparameter is None:
            latticeparameter = (1.0, 1.0, 1.0, 90.0, 90.0, 90.0)
        if polefamilies is None:
            polefamilies = default
    else:
        if spacegroup is None:
            return addphase(libtype="FCC", phasename=phasename, latticeparameter=latticeparameter,
                            polefamilies=polefamilies, angTol=angTol)
        if latticeparameter is None:
            latticeparameter = (1.0, 1.0, 1.0, 90.0, 90.0, 90.0)
        if polefamilies is None:
            polefamilies = FCC_POLES
    return BandIndexer(phasename=phasename, spacegroup=spacegroup, latticeparameter=latticeparameter,
                       polefamilies=np.atleast_2d(polefamilies), angTol=angTol, nband_earlyexit=nband_earlyexit,
                       pointgroup=pointgroup, pointgroupid=pointgroupid)


class BandIndexer:
    """A phase's triplet library and its band indexing on the device (tripletvote.py:171-
    1018): the library on the host, the vote, band assignment, QUEST orientation and
    confidence metric on the device."""

    def __init__(self, phasename=None, spacegroup=225, latticeparameter=(1.0, 1.0, 1.0, 90.0, 90.0, 90.0),
                 polefamilies=FCC_POLES, angTol: float = 2.0, nband_earlyexit: int = 8, pointgroup=" ",
                 pointgroupid=None):
        self.phasename = "" if phasename is None else str(phasename)
        self.spacegroup = int(spacegroup)
        self.latticeparameter = np.array(latticeparameter, dtype=np.float64)
        self.angTol = float(angTol)
        self.nband_earlyexit = int(nband_earlyexit)
        self.pointgroup = pointgroup
        self.pointgroupid = pointgroupid
        self.high_fidelity = True
        self.simpleqweights = False
        self.lut = triplet_lut()
        lib = triplet_library(polefamilies, self.spacegroup, self.latticeparameter)
        self.lauecode = lib["lauecode"]
        self.crystalmats = lib["crystal"]
        self.polefamilies = lib["polefamilies"]
        self.npolefamilies = lib["nfam"]
        self.angpairs = lib["angpairs"]
        self.angtriplets = lib["angtriplets"]
        self.completelib = lib["completelib"]
        self.families = lib["families"]
        self._dev = {}

    def __getstate__(self):
        """Pickled (Ray's object store pickles the indexer with its phases,
        _ebsd_index_parallel.py:396) without the device copies of the library, which are
        made again on first use in the process that unpickles."""
        state = self.__dict__.copy()
        state["_dev"] = {}
        return state

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

    def _device_complete(self, device):
        """The complete pole list on the device, for ebsd_bandindex (made once per device)."""
        import torch
        key = ("complete", str(device))
        if key not in self._dev:
            c = self.completelib
            t = lambda x, dt: torch.as_tensor(np.ascontiguousarray(x, dtype=dt), device=device)
            self._dev[key] = (t(c["polesCart"], np.float32), t(c["polesCart"], np.float64),
                              t(c["familyid"], np.int32), t(c["angTable"][c["famIndex"]], np.float64),
                              t(c["famIndex"], np.int32), t(c["nFamily"], np.int32))
        return self._dev[key]

    def bandindex(self, band_norms, band_intensity=None, band_widths=None, verbose=0, stream=None,
                  return_assignment=False):
        """BandIndexer.bandindex (tripletvote.py:492-633) on the device: the vote, band
        assignment, QUEST orientation and confidence metric.  ``band_norms`` (npats, nbands,
        3) or (nbands, 3), numpy or a device tensor; ``band_intensity`` likewise (default
        ones).  Returns (avequat, fit, cm2, polematch, nMatch, ij, acc_correct, bandfit) as the
        reference does -- numpy arrays for host input, device tensors for device input --
        and, with ``return_assignment``, the assignment's own (polevalid, fit, per-band
        fit) after them."""
        import torch
        if not self.high_fidelity:
            raise NotImplementedError("high_fidelity=False (the TRIAD rotation as the orientation) is not built")
        host = not (isinstance(band_norms, torch.Tensor) and band_norms.is_cuda)
        dev = torch.device("cuda", torch.cuda.current_device()) if host else band_norms.device
        bn = torch.as_tensor(np.asarray(band_norms, dtype=np.float32) if host else band_norms, device=dev)
        bn = (bn.unsqueeze(0) if bn.dim() == 2 else bn).to(torch.float32).contiguous()
        npats, nb = int(bn.shape[0]), int(bn.shape[1])
        if band_intensity is None:
            inten = torch.ones((npats, nb), dtype=torch.float32, device=dev)
        else:
            inten = torch.as_tensor(band_intensity, device=dev).to(torch.float32).reshape(npats, nb).contiguous()
        accW, fam, rank, cm, acc = self.tripvote(bn, inten, stream=stream)
        pf, pc64, famid, angrows, famidx, famcnt = self._device_complete(dev)
        f64 = dict(dtype=torch.float64, device=dev)
        quat = torch.empty((npats, 4), **f64)
        fit = torch.empty(npats, **f64)
        bandfit = torch.empty((npats, nb), **f64)
        cm2 = torch.empty(npats, dtype=torch.float32, device=dev)
        pm = torch.empty((npats, nb), dtype=torch.int64, device=dev)
        nmatch = torch.empty(npats, dtype=torch.int64, device=dev)
        ij = torch.empty((npats, 4), dtype=torch.int64, device=dev)
        accc = torch.empty(npats, dtype=torch.int32, device=dev)
        pv = torch.empty((npats, nb), dtype=torch.uint8, device=dev)
        afit = torch.empty(npats, dtype=torch.float32, device=dev)
        fitb = torch.empty((npats, nb), dtype=torch.float32, device=dev)
        from ._hip import keep_for_stream, launch_stream
        s = launch_stream(stream, dev)
        lib = load_tripvote()
        P = ctypes.c_void_p
        ptr = lambda t: P(t.data_ptr())
        rc = lib.ebsd_bandindex(ptr(bn), ptr(inten), ctypes.c_int64(npats), nb, ptr(accW), ptr(fam), ptr(rank),
                                ptr(acc), int(self.npolefamilies), ptr(pf), ptr(pc64), ptr(famid), int(pf.shape[0]),
                                ptr(angrows), ptr(famidx), ptr(famcnt), ctypes.c_double(self.angTol),
                                int(self.nband_earlyexit), 6, int(bool(self.simpleqweights)), ptr(quat), ptr(fit),
                                ptr(cm2), ptr(pm), ptr(nmatch), ptr(ij), ptr(accc), ptr(bandfit), ptr(pv),
                                ptr(afit), ptr(fitb), P(s.cuda_stream))
        check_tripvote(lib, rc)
        outs = [quat, fit, cm2, pm, nmatch, ij, accc, bandfit]
        if return_assignment:
            outs += [pv, afit, fitb]
        keep_for_stream(s, bn, inten, accW, fam, rank, cm, acc, *outs)
        if host:
            if stream is not None:
                s.synchronize()
            return tuple(o.cpu().numpy() for o in outs)
        return tuple(outs)


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
           "laue_class", "laue_rotations", "Crystal", "hex4_to_hex3", "hex3_to_hex4", "HEX_POLES",
           "FCC_POLES", "BCC_POLES"]
