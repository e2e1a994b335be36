"""The triplet-vote oracle (oracle/tripvote_ref.py, csrc/ref_tripvote.c) and the cubic
triplet library (pyebsdindex_amd/tripletvote.py), pinned on the CPU:

* the library's pair-angle counts are the reference's own test values
  (pyebsdindex/tests/test_tripletvote.py:32-47) and its angles are the closed-form
  angles between cubic poles; the triplet table equals the reference's
  (tripletvote.py:207-213);
* a hand-solved scalene triangle: which band takes which family follows from the
  geometry (the vertex between the 30 and 40 degree edges is family 0, ...), whatever
  order the bands come in;
* an invalid last band leaves the other bands' votes bit-identical to those of the
  shorter band list (the reference's extra votes for near-equal angles swap bands by
  their place in the triplet, so a permutation of the bands is not a symmetry of the
  vote and is not tested);
* on exact cubic poles in random orientations the vote names the true family of most
  bands (a wrong vertex order or unsort would scramble them).
"""
import itertools

import numpy as np
import pytest

from oracle import tripvote_ref
from pyebsdindex_amd import tripletvote as tv

from helpers import cubic_bands


def _vote(bn, ind, inten=None, tol=None):
    return tripvote_ref.tripvote(bn, inten, ind.lut, ind.angTol if tol is None else tol,
                                 ind.angtriplets["angles"], ind.angtriplets["familyid"], ind.npolefamilies)


def test_lut_is_the_reference_table():
    # tripletvote.py:207-213: LUT[:, luta[i]] = lutb[i]
    luta = [[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]]
    lutb = [[0, 1, 2], [1, 0, 2], [0, 2, 1], [2, 0, 1], [1, 2, 0], [2, 1, 0]]
    want = np.zeros((3, 3, 3, 3), dtype=np.int64)
    for a, b in zip(luta, lutb):
        want[:, a[0], a[1], a[2]] = b
    assert np.array_equal(tv.triplet_lut(), want)


@pytest.mark.parametrize("name,poles,nang,nuniq", [("FCC", tv.FCC_POLES, 21, 17), ("BCC", tv.BCC_POLES, 34, 28)])
def test_library_golden_counts(name, poles, nang, nuniq):
    ind = tv.addphase(name)
    assert np.array_equal(ind.polefamilies, poles)
    a = ind.angpairs["angles"]
    assert a.size == nang and np.unique(a).size == nuniq


def test_library_golden_counts_hcp_and_triclinic():
    """The reference's own values (pyebsdindex/tests/test_tripletvote.py:50-96): HCP (space
    group 194, c/a 1.63, the eight default four-index families) 82 pair angles, 74 distinct;
    a triclinic cell (2, 3, 4, 70, 100, 120) with the 26 poles of {-1,0,1}^3 -> 13 families,
    78 angles, 77 distinct."""
    import itertools
    ph = tv.addphase("HCP")
    assert np.allclose(ph.polefamilies, [[1, 0, -1, 0], [0, 0, 0, 2], [1, 0, -1, 1], [1, 0, -1, 2],
                                         [1, 1, -2, 0], [1, 0, -1, 3], [1, 1, -2, 2], [2, 0, -2, 1]])
    a = ph.angpairs["angles"]
    assert a.size == 82 and np.unique(a).size == 74
    hkl = np.asarray(list(itertools.product(*[np.arange(-1, 2)] * 3)), dtype=int)
    hkl = hkl[~np.all(hkl == 0, axis=1)]
    tri = tv.addphase(spacegroup=1, latticeparameter=[2, 3, 4, 70, 100, 120], nband_earlyexit=5, polefamilies=hkl)
    assert tri.spacegroup == 1 and tri.nband_earlyexit == 5
    assert np.allclose(tri.latticeparameter, [2, 3, 4, 70, 100, 120])
    assert tri.npolefamilies == hkl.shape[0] // 2
    for h in tri.polefamilies:
        assert any(np.array_equal(h, x) for x in hkl)
    a = tri.angpairs["angles"]
    assert a.size == 78 and np.unique(a).size == 77


@pytest.mark.parametrize("phase", ["FCC", "BCC", "HCP", "triclinic"])
def test_library_angles_clear_of_rounding_edges(phase):
    """The families here are exact integer orbits; the reference rotates poles with float32
    quaternions, so its pole indices carry ~1e-7 of noise.  Pair angles are rounded to 0.01
    degree, so the two agree unless an exact angle lies within that noise of a rounding
    edge: every exact pair angle (in units of 0.01 degree) is at least 1e-3 from an edge."""
    import itertools
    if phase == "triclinic":
        hkl = np.asarray(list(itertools.product(*[np.arange(-1, 2)] * 3)), dtype=int)
        ph = tv.addphase(spacegroup=1, latticeparameter=[2, 3, 4, 70, 100, 120], polefamilies=hkl[~np.all(hkl == 0, 1)])
    else:
        ph = tv.addphase(phase)
    lib = tv.triplet_library(ph.polefamilies, ph.spacegroup, ph.latticeparameter)
    rmt = ph.crystalmats.reciprocalMetricTensor
    poles = tv.hex4_to_hex3(ph.polefamilies) if ph.lauecode in (6, 62) else ph.polefamilies
    for i in range(lib["nfam"]):
        for j in range(i, lib["nfam"]):
            x = np.degrees(np.arccos(np.abs(tv._pole_cos(poles[i], lib["families"][j], rmt).ravel()))) * 100
            frac = np.abs(x - np.floor(x) - 0.5)
            assert frac.min() > 1e-3, (phase, i, j, x[np.argmin(frac)])


def test_hcp_families_are_the_hexagonal_orbits():
    """6/mmm orbits of the HCP families (both signs of a plane normal kept, as the
    reference's complete list does): {10-10} and {11-20} 6 normals, {0002} 2, {10-11} 12;
    every variant of a family makes the family's own angle with c, and all are unit."""
    ph = tv.addphase("HCP")
    c = ph.completelib
    n = dict(zip(range(8), c["nFamily"]))
    assert n[0] == 6 and n[1] == 2 and n[2] == 12 and n[4] == 6
    z = np.array([0.0, 0.0, 1.0])
    for f in range(8):
        sel = c["familyid"] == f
        cz = np.abs(c["polesCart"][sel] @ z)
        assert np.allclose(cz, cz[0], atol=1e-12)
    assert np.allclose(np.linalg.norm(c["polesCart"], axis=1), 1.0)


def test_library_closed_form_angles():
    ind = tv.addphase("FCC")
    a, f = ind.angpairs["angles"], ind.angpairs["familyid"]
    def angs(i, j):
        return set(np.round(a[(f[:, 0] == i) & (f[:, 1] == j)], 2))
    deg = lambda c: round(float(np.degrees(np.arccos(c))), 2)
    # {002}: 90;  {002}-{111}: 54.74;  {111}-{111}: 70.53;  {002}-{022}: 45, 90
    assert angs(0, 0) == {90.0}
    assert angs(0, 1) == {deg(1 / np.sqrt(3))}
    assert angs(1, 1) == {deg(1 / 3)}
    assert angs(0, 2) == {45.0, 90.0}
    # {002}-{113}: cos = 3/sqrt(11), 1/sqrt(11)
    assert angs(0, 3) == {deg(3 / np.sqrt(11)), deg(1 / np.sqrt(11))}
    fams = ind.families
    assert [f_.shape[0] for f_ in fams] == [6, 8, 12, 24]
    # every triplet is ascending, its ids are families, its angles are pair angles
    t, ti = ind.angtriplets["angles"], ind.angtriplets["familyid"]
    assert np.all(np.diff(t, axis=1) >= 0) and ti.min() >= 0 and ti.max() < ind.npolefamilies
    assert set(np.round(t.ravel(), 2)) <= set(np.round(a, 2))


def _triangle(ab, ac, bc):
    """Unit vectors a, b, c with the given pairwise angles (degrees)."""
    r = np.radians
    a = np.array([0.0, 0.0, 1.0])
    b = np.array([np.sin(r(ab)), 0.0, np.cos(r(ab))])
    cz = np.cos(r(ac))
    cx = (np.cos(r(bc)) - cz * np.cos(r(ab))) / np.sin(r(ab))
    c = np.array([cx, np.sqrt(1 - cx * cx - cz * cz), cz])
    return np.stack([a, b, c]).astype(np.float32)


def test_hand_solved_triangle():
    """Library row (30, 40, 50) with families (0, 1, 2) at vertices (A, B, C): A joins the
    30 and 40 degree edges, B the 30 and 50, C the 40 and 50.  Bands with ab = 30,
    ac = 50, bc = 40: a joins 30 and 50 -> family 1, b joins 30 and 40 -> family 0,
    c joins 40 and 50 -> family 2; one vote each, weight angTol for exact angles."""
    lut = tv.triplet_lut()
    lib_a = np.array([[30.0, 40.0, 50.0]])
    lib_i = np.array([[0, 1, 2]])
    bands = _triangle(30, 50, 40)
    want = {0: 1, 1: 0, 2: 2}
    for order in itertools.permutations(range(3)):
        bn = bands[list(order)]
        accW, fam, rank, cm, acc = tripvote_ref.tripvote(bn, None, lut, 2.0, lib_a, lib_i, 3)
        for pos, band in enumerate(order):
            assert fam[0, pos] == want[band], (order, pos)
            col = acc[0, :, pos]
            assert col[want[band]] == 1 and col.sum() == 1
            assert abs(accW[0, want[band], pos] - 2.0) < 1e-4
            # one vote of weight ~2: trunc(sum) = 1, cm = 2 - 0, rank = (3 - pos)/3 * cm * 1
            assert abs(cm[0, pos] - accW[0, want[band], pos]) < 1e-6
            assert abs(rank[0, pos] - (3 - pos) / 3 * cm[0, pos]) < 1e-6


def test_triangle_outside_tolerance_gets_no_vote():
    lut = tv.triplet_lut()
    accW, fam, rank, cm, acc = tripvote_ref.tripvote(_triangle(30, 50, 40), None, lut, 2.0,
                                                     np.array([[30.0, 40.0, 52.5]]), np.array([[0, 1, 2]]), 3)
    assert acc.sum() == 0 and np.all(cm == 0) and np.all(rank == 0) and np.all(fam == 0)


@pytest.mark.parametrize("phase", ["FCC", "BCC"])
def test_invalid_last_band_is_the_shorter_list(phase):
    ind = tv.addphase(phase)
    bn, _ = cubic_bands(ind, 40, 9, seed=3, noise_deg=0.2)
    inten = np.ones((40, 9), np.float32)
    inten[:, 8] = 0.0
    full = _vote(bn, ind, inten)
    short = _vote(np.ascontiguousarray(bn[:, :8]), ind)
    assert np.array_equal(full[0][:, :, :8], short[0])      # weighted votes, bit for bit
    assert np.array_equal(full[4][:, :, :8], short[4])      # counts
    assert np.all(full[4][:, :, 8] == 0)
    assert np.array_equal(full[1][:, :8], short[1])
    assert np.array_equal(full[3][:, :8], short[3])


@pytest.mark.parametrize("phase", ["FCC", "BCC"])
def test_exact_poles_name_their_family(phase):
    ind = tv.addphase(phase)
    bn, fam_true = cubic_bands(ind, 200, 9, seed=11, noise_deg=0.0)
    accW, fam, rank, cm, acc = _vote(bn, ind)
    agree = float(np.mean(fam == fam_true))
    assert agree > 0.85, agree
    # the true family of a band has at least one vote in most of its 28 triplets
    got = np.take_along_axis(acc, fam_true[:, None, :], axis=1)[:, 0, :]
    assert np.median(got) >= 20


@pytest.mark.parametrize("phase,tol", [("FCC", 2.0), ("BCC", 0.5), ("BCC", 3.0)])
def test_row_index_lists_every_candidate(phase, tol):
    """The device's row index (tripletvote.row_index) lists, for each bin, every library
    row whose first angle is within tol of an x in the bin -- checked at the bin edges and
    inside -- in ascending order."""
    ind = tv.addphase(phase)
    a0 = ind.angtriplets["angles"][:, 0]
    start, rows, bw, built = tv.row_index(ind.angtriplets["angles"], tol)
    assert built == tol
    nb = start.size - 1
    for b in range(nb):
        lst = rows[start[b]:start[b + 1]]
        assert np.all(np.diff(lst) > 0)
        xs = [b * bw, b * bw + bw / 2, np.nextafter((b + 1) * bw, 0)]
        if b == nb - 1:
            xs += [95.0, 180.0]
        if b == 0:
            xs += [-1.0]
        for x in np.float32(xs):
            need = np.nonzero(np.abs(a0 - np.float64(x)) <= tol)[0]
            assert set(need) <= set(lst), (b, x)
            # the kernel's float32 bin of x may be off by one next to an edge: the
            # neighbouring bin's list must hold the rows too
            kb = int(np.floor(np.float32(x) * np.float32(1.0 / bw)))
            kb = min(max(kb, 0), nb - 1)
            assert set(need) <= set(rows[start[kb]:start[kb + 1]]), (b, x, kb)


def _angle_deg(x, y, fused):
    """The float32 inter-band angle of the oracle (((x0 y0) + (x1 y1)) + (x2 y2), no
    contraction) or of a fused form (fma(x2, y2, fma(x1, y1, x0 y0)), each fma rounded once
    to float32); arccos correctly rounded to float32, then times RADEG in float64."""
    from fractions import Fraction
    x, y = np.asarray(x, np.float32), np.asarray(y, np.float32)
    if fused:
        acc = np.float32(x[0] * y[0])
        for k in (1, 2):
            acc = np.float32(float(Fraction(float(x[k])) * Fraction(float(y[k])) + Fraction(float(acc))))
        c = acc
    else:
        c = np.float32(np.float32(x[0] * y[0]) + np.float32(x[1] * y[1])) + np.float32(x[2] * y[2])
    c = min(abs(np.float32(c)), np.float32(1.0))
    return np.float32(float(np.float32(np.arccos(np.float64(c)))) * (180.0 / np.pi))


def test_cosine_form_at_the_tolerance_edge():
    """Parity unpinned, made visible: the reference's inter-band cosines come from a float32
    2-D ``.dot`` (tripletvote.py:1192), which Numba hands to BLAS; whether that sgemm fuses
    the multiply-adds is not known here.  The oracle and the device evaluate
    ((x0 y0) + (x1 y1)) + (x2 y2) without contraction.  For a band pair whose angle differs
    by an ulp between the two forms, a library row placed exactly angTol away from the
    oracle's angle gets the oracle's vote, and would not get it from the fused form."""
    rng = np.random.default_rng(17)
    a = b = None
    for _ in range(200000):
        x = rng.normal(size=3)
        y = x + rng.normal(scale=0.9, size=3)
        x = (x / np.linalg.norm(x)).astype(np.float32)
        y = (y / np.linalg.norm(y)).astype(np.float32)
        ga, gf = _angle_deg(x, y, False), _angle_deg(x, y, True)
        if ga != gf and 15.0 < ga < 60.0:
            a, b = x, y
            break
    assert a is not None, "no pair with differing cosine forms found"
    # a third band far (>= 10 degrees) from the pair's angle on both edges
    for _ in range(10000):
        z = rng.normal(size=3)
        z = (z / np.linalg.norm(z)).astype(np.float32)
        e1, e2 = _angle_deg(a, z, False), _angle_deg(b, z, False)
        if min(abs(e1 - ga), abs(e2 - ga), abs(e1 - e2)) > 10.0:
            break
    bn = np.stack([a, b, z])[None]
    angs = np.array([ga, e1, e2], dtype=np.float32)  # (AB, AC, BC)
    srt = np.argsort(angs)
    # the library row: the pattern's own angles, the pair's angle moved by exactly
    # angTol = 1 away from the oracle's value, on the side the fused value lies away from
    step = 1.0 if gf < ga else -1.0
    row = angs.astype(np.float64)
    row[0] = float(ga) + step
    lib = row[srt][None]
    fam = np.array([[0, 0, 0]], dtype=np.int64)
    accW, bandFam, rank, cm, acc = tripvote_ref.tripvote(bn, None, tv.triplet_lut(), 1.0, lib, fam, 1)
    assert acc.sum() >= 3            # the oracle's angle is angTol away: the row votes
    assert abs(row[0] - float(gf)) > 1.0   # the fused form's angle is just outside angTol
    # one ulp further the oracle too leaves the row out
    row2 = row.copy()
    row2[0] = np.nextafter(row[0], row[0] + step)
    accW, bandFam, rank, cm, acc = tripvote_ref.tripvote(bn, None, tv.triplet_lut(), 1.0, row2[srt][None], fam, 1)
    assert acc.sum() == 0
