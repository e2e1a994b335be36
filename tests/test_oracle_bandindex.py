"""The band-indexing oracle after the vote (oracle/tripvote_ref.py, csrc/ref_bandindex.c):
band assignment (_assign_bands_nb2), QUEST weights and fit, the confidence metric, pinned on
the CPU:

* exact poles of known rotations (FCC, BCC, HCP): every band is assigned to a pole of its
  true family, the QUEST quaternion rotates each band normal onto its pole, and the
  rotation equals the true one up to the phase's Laue-class symmetry;
* QUEST solves Wahba's problem: on noisy weighted vector pairs its rotation equals the SVD
  solution (a library routine) to 1e-9;
* the QUEST weights keep the six best-scoring matched bands with their intensities,
  normalised; the confidence metric equals a hand-worked value;
* noisy bands (0.3 degree) still give the true orientation within half a degree.
"""
import numpy as np
import pytest

from helpers import phase_bands, quat_to_rot, symmetric_misorientation_deg
from oracle import tripvote_ref as tr
from oracle.orientation import wahba_svd
from pyebsdindex_amd import tripletvote as tv


@pytest.mark.parametrize("phase,nb", [("FCC", 9), ("BCC", 9), ("HCP", 10)])
def test_exact_poles_recover_rotation(phase, nb):
    ind = tv.addphase(phase)
    bn, pole, fam, Rt = phase_bands(ind, 40, nb, seed=3)
    q, fit, cm2, pm, nm, ij, accc, bandfit = tr.bandindex(ind, bn)
    famid = ind.completelib["familyid"]
    rots = tv.laue_rotations(ind.lauecode)
    assert (nm >= min(nb, ind.nband_earlyexit)).all(), nm
    # fit = arccos(mean dot): near 1 the float32 rounding of the normals (~3e-8) shows as
    # about 0.01 degree
    assert (fit < 0.03).all(), fit.max()
    for p in range(bn.shape[0]):
        Rq = quat_to_rot(q[p])
        assert abs(np.linalg.det(Rq) - 1.0) < 1e-9
        good = pm[p] >= 0
        # the rotation takes every assigned band normal onto its library pole
        onto = bn[p][good].astype(np.float64) @ Rq.T
        assert np.abs(onto - ind.completelib["polesCart"][pm[p][good]]).max() < 1e-4
        # and it is the true rotation, up to the Laue-class symmetry
        assert symmetric_misorientation_deg(Rq.T, Rt[p], rots) < 1e-3
    # exact poles: each matched band names its true family
    ok = (pm >= 0) & (famid[pm.clip(0)] == fam)
    assert ok.mean() == 1.0
    # acc_correct counts the matched bands whose pole family is the vote's family for them:
    # the vote names most families right on exact poles
    assert (accc <= (pm >= 0).sum(axis=1)).all() and accc.mean() >= 0.8 * nb
    assert ((cm2 > 0) & (cm2 <= 1)).all()


def test_quest_solves_wahba():
    rng = np.random.default_rng(9)
    from scipy.spatial.transform import Rotation
    npats, nb = 64, 9
    R = Rotation.random(npats, random_state=rng).as_matrix()
    poles = rng.normal(size=(npats, nb, 3))
    poles /= np.linalg.norm(poles, axis=2, keepdims=True)
    bands = np.einsum("pij,pkj->pki", np.transpose(R, (0, 2, 1)), poles)   # R^T c
    bands += rng.normal(scale=0.01, size=bands.shape)
    bands /= np.linalg.norm(bands, axis=2, keepdims=True)
    w = rng.uniform(0.1, 1.0, (npats, nb))
    w[:, 7:] = 0.0                                                         # two bands left out
    wn = w / w.sum(axis=1, keepdims=True)
    q, lam, mad, bd = tr.quest(poles, bands, wn)
    for p in range(npats):
        Rq = quat_to_rot(q[p])
        Rs = wahba_svd(poles[p], bands[p], wn[p])
        assert np.abs(Rq - Rs).max() < 1e-9, p
        # lambda is the largest eigenvalue of Davenport's K: the loss is 1 - lambda
        loss = np.sum(wn[p] * np.sum((poles[p] - bands[p] @ Rs.T) ** 2, axis=1)) / 2.0
        assert abs((1.0 - lam[p]) - loss) < 1e-9
        assert np.allclose(bd[p, :7], np.sum((bands[p, :7] @ Rq.T) * poles[p, :7], axis=1))
        assert np.all(bd[p, 7:] == -1.0)


def test_quest_weights_and_cm_by_hand():
    # two families, four bands: votes (weighted) and counts per (family, band)
    accW = np.array([[[6.0, 0.5, 3.0, 1.0], [1.0, 4.0, 0.0, 2.0]]], np.float32)
    acc = np.array([[[3.0, 1.0, 2.0, 1.0], [1.0, 2.0, 1.0, 4.0]]], np.float32)
    famid = np.array([0, 0, 1, 1], np.int32)          # library poles 0,1 -> family 0; 2,3 -> 1
    pm = np.array([[0, 2, 1, 3]], np.int64)           # bands -> families 0, 1, 0, 1
    pv = np.array([[1, 1, 1, 1]], np.uint8)
    inten = np.array([[10.0, 20.0, 30.0, 40.0]], np.float32)
    w = tr.quest_weights(famid, accW, acc, pm, pv, inten, nfit=3)
    # scores: 6/3 = 2, 4/2 = 2, 3/2 = 1.5, 2/4 = 0.5 -> the top three by score, ties in
    # reverse index order (a flipped stable sort): bands 1, 0, 2
    want = np.array([10.0, 20.0, 30.0, 0.0]) / 60.0
    assert np.allclose(w[0], want, atol=1e-7)
    cm = tr.calc_cm(accW, pm, pv, famid)
    # matched votes 6 + 4 + 3 + 2 = 15 over the clipped total 6+1+3+1+1+4+1+2 = 19
    assert cm[0] == np.float32(15.0) / np.float32(19.0)
    pv2 = np.array([[1, 1, 0, 0]], np.uint8)
    assert tr.calc_cm(accW, pm, pv2, famid)[0] == -1.0   # fewer than three matched bands


@pytest.mark.parametrize("phase", ["FCC", "HCP"])
def test_noisy_bands_orientation(phase):
    ind = tv.addphase(phase)
    bn, pole, fam, Rt = phase_bands(ind, 60, 9, seed=11, noise_deg=0.3)
    q, fit, cm2, pm, nm, ij, accc, bandfit = tr.bandindex(ind, bn)
    rots = tv.laue_rotations(ind.lauecode)
    errs = [symmetric_misorientation_deg(quat_to_rot(q[p]).T, Rt[p], rots) for p in range(bn.shape[0])]
    assert np.median(errs) < 0.3 and np.max(errs) < 1.0, (np.median(errs), np.max(errs))
    assert (fit < 1.0).all()
