"""Shared test helpers: parameter sets, oracle runs and near-tie-aware band comparison."""
import numpy as np

from oracle import fp64, ref_cpu

# tSigma=1, rSigma=2 give the 9x5 kernel, peakPad (19, 11) and padding (19, 11) of the workloads
STD = dict(nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9)


def make_pair(shape, **kw):
    """(product BandDetect, oracle RefBandDetect) with identical parameters."""
    from pyebsdindex_amd import BandDetect
    p = dict(STD)
    p.update(kw)
    bd = BandDetect(patDim=shape, **p)
    ref = ref_cpu.RefBandDetect(shape, **{k: v for k, v in p.items() if k != 'background_mode'})
    return bd, ref


def oracle64(ref, pats, background=None):
    return fp64.find_bands64(pats, ref.radonPlan.indexPlan, ref.kernel[0], ref.rdnmask, ref.padding,
                             ref.peakPad, ref.nBands, ref.radonPlan.theta, ref.radonPlan.rho,
                             background=background)


def eps_local_max(C, r, c, hr, ht, eps):
    """True if C[r,c] is within eps of the clamped-window maximum (a local max up to eps)."""
    r0, r1 = max(r - hr, 0), min(r + hr, C.shape[0] - 1)
    c0, c1 = max(c - ht, 0), min(c + ht, C.shape[1] - 1)
    return C[r, c] >= C[r0:r1 + 1, c0:c1 + 1].max() - eps


def compare_bands(got, res64, C64, padding, peakPad, rel_eps=2e-5, subpix_tol=1e-3):
    """Compare device band output with the FP64 oracle.

    Integer peak bins must be identical, except where the FP64 values of the two choices
    differ by less than ``rel_eps * max|C|`` (an operational exact tie: the float32 paths
    cannot order them).  Returns (n_slots_compared, n_tie_slots).  Sub-pixel positions of
    matching slots must agree within ``subpix_tol`` bins."""
    p0, p1 = int(padding[0]), int(padding[1])
    hr, ht = int(peakPad[0]) // 2, int(peakPad[1]) // 2
    N, nB = res64['max'].shape
    ties = 0
    for q in range(N):
        C = C64[q]
        eps = rel_eps * np.abs(C).max()
        for i in range(nB):
            gv, ov = int(got['valid'][q, i]), int(res64['valid'][q, i])
            gl = got['maxloc'][q, i].astype(np.int64)
            ol = res64['maxloc'][q, i].astype(np.int64)
            if gv == ov == 1 and np.array_equal(gl, ol):
                d = np.abs(got['aveloc'][q, i] - res64['aveloc'][q, i]).max()
                assert d <= subpix_tol, (q, i, got['aveloc'][q, i], res64['aveloc'][q, i])
                continue
            if gv == 0 and ov == 0:
                continue
            # disagreement: must be an exact tie up to eps
            gval = C[gl[0] + p0, gl[1] + p1] if gv else 0.0
            oval = C[ol[0] + p0, ol[1] + p1] if ov else 0.0
            ok = abs(gval - oval) <= eps
            if gv:
                ok = ok and eps_local_max(C, gl[0] + p0, gl[1] + p1, hr, ht, eps)
            assert ok, (f"pattern {q} slot {i}: device {gl} valid {gv} C={gval}, "
                        f"oracle {ol} valid {ov} C={oval}, eps={eps}")
            ties += 1
    return N * nB, ties


def orientation_agreement(got, res, meta, pc, shape):
    """Fit an orientation to each pattern's bands from ``got`` and from ``res`` (structured
    band records with theta, rho, valid, max) with the same assignment to the generator's
    planes and the same weighted Wahba fit (oracle/orientation.py), on the bands both found
    at the same plane.  Returns (patterns fitted, worst got-vs-res misorientation in deg,
    median got-vs-true misorientation in deg)."""
    from oracle import gnomonic
    from oracle import orientation as ori
    from pyebsdindex_amd.synthetic import quat_to_matrix
    n_got = gnomonic.radon2pole(got['theta'], got['rho'], pc, shape)
    n_res = gnomonic.radon2pole(res['theta'], res['rho'], pc, shape)
    R_true = quat_to_matrix(meta['quat'])
    worst, fitted, vs_true = 0.0, 0, []
    for q in range(n_got.shape[0]):
        a = ori.assign_bands(n_got[q], meta['g_det'][q], meta['normals'], got['valid'][q])
        b = ori.assign_bands(n_res[q], meta['g_det'][q], meta['normals'], res['valid'][q])
        # a band the two paths pick differently (near-tie peaks) changes the data set
        common = np.intersect1d(a[2], b[2])
        ia, ib = np.searchsorted(a[2], common), np.searchsorted(b[2], common)
        same = np.all(a[1][ia] == b[1][ib], axis=1)
        ia, ib = ia[same], ib[same]
        if ia.size < 3:
            continue
        Ra = ori.wahba_svd(a[0][ia], a[1][ia], got['max'][q][a[2][ia]].astype(np.float64))
        Rb = ori.wahba_svd(b[0][ib], b[1][ib], res['max'][q][b[2][ib]].astype(np.float64))
        worst = max(worst, ori.misorientation_deg(Ra, Rb))
        vs_true.append(ori.misorientation_deg(Ra, R_true[q]))
        fitted += 1
    return fitted, worst, float(np.median(vs_true)) if vs_true else float('inf')


def cubic_bands(indexer, npats, nbands, seed=0, noise_deg=0.0):
    """Band normals of npats random orientations of a cubic phase: nbands distinct poles
    (no two within 3 degrees, antiparallel counted as the same) drawn from the indexer's
    pole families, rotated, optionally tilted by Gaussian noise of noise_deg degrees per
    component; float32 (npats, nbands, 3) and the true family of each band."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(seed)
    fams = indexer.families
    allp = np.concatenate([f / np.linalg.norm(f, axis=1, keepdims=True) for f in fams])
    famid = np.concatenate([np.full(f.shape[0], i) for i, f in enumerate(fams)])
    out = np.zeros((npats, nbands, 3), np.float32)
    fam = np.zeros((npats, nbands), np.int32)
    cos3 = np.cos(np.radians(3.0))
    for p in range(npats):
        chosen = []
        for idx in rng.permutation(allp.shape[0]):
            if all(abs(allp[idx] @ allp[c]) < cos3 for c in chosen):
                chosen.append(idx)
            if len(chosen) == nbands:
                break
        R = Rotation.random(random_state=rng).as_matrix()
        n = allp[chosen] @ R.T
        if noise_deg:
            n = n + rng.normal(0.0, np.radians(noise_deg), n.shape)
            n /= np.linalg.norm(n, axis=1, keepdims=True)
        out[p, :len(chosen)] = n
        fam[p, :len(chosen)] = famid[chosen]
    return out, fam


def quat_to_rot(q):
    """Active rotation matrix of a unit quaternion (w, x, y, z): R v = q v q* -- the rotation
    the reference's QUEST applies to band normals (quat_vector with p = +1)."""
    w, x, y, z = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def phase_bands(indexer, npats, nbands, seed=0, noise_deg=0.0, min_sep=3.0):
    """Band normals of npats random orientations of any phase: nbands poles of its complete
    pole list (Cartesian, with its lattice), no two within min_sep degrees (antiparallel
    counted as the same), rotated into the detector frame by R_true, optionally tilted by
    Gaussian noise of noise_deg degrees per component.  Returns float32 normals (npats,
    nbands, 3), the true pole index and family (npats, nbands) and R_true (npats, 3, 3)."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(seed)
    lib = indexer.completelib
    allp = np.asarray(lib["polesCart"], np.float64)
    famid = np.asarray(lib["familyid"])
    out = np.zeros((npats, nbands, 3), np.float32)
    pole = np.full((npats, nbands), -1, np.int64)
    fam = np.full((npats, nbands), -1, np.int32)
    Rs = np.zeros((npats, 3, 3))
    cs = np.cos(np.radians(min_sep))
    for p in range(npats):
        chosen = []
        for idx in rng.permutation(allp.shape[0]):
            if all(abs(allp[idx] @ allp[c]) < cs for c in chosen):
                chosen.append(idx)
            if len(chosen) == nbands:
                break
        R = Rotation.random(random_state=rng).as_matrix()
        n = allp[chosen] @ R.T
        if noise_deg:
            n = n + rng.normal(0.0, np.radians(noise_deg), n.shape)
            n /= np.linalg.norm(n, axis=1, keepdims=True)
        out[p, :len(chosen)] = n
        pole[p, :len(chosen)] = chosen
        fam[p, :len(chosen)] = famid[chosen]
        Rs[p] = R
    return out, pole, fam, Rs


def symmetric_misorientation_deg(Ra, Rb, rots):
    """Smallest rotation angle (degrees) between Ra and Rb over the crystal symmetry rotations
    ``rots`` (Cartesian, (n, 3, 3)) applied on the crystal side: min_S angle(Ra^T Rb S)."""
    from oracle.orientation import misorientation_deg
    return min(misorientation_deg(Ra, Rb @ S) for S in rots)


def adj_intensity(bands, rhomax):
    """The band weights EBSDIndexer passes to bandindex (_ebsd_index_single.py:757-763):
    max scaled down away from rho = 0 and near theta = 0 or 180 degrees, zero for invalid
    bands (float32)."""
    rho = np.asarray(bands["rho"], np.float32)
    theta = np.asarray(bands["theta"], np.float32)
    rm = np.float32(rhomax)
    a = (np.float32(-1) * np.abs(rho) * np.float32(0.5) / rm + np.float32(1)) * np.asarray(bands["max"], np.float32)
    a *= ((theta > (2 * np.pi / 180)).astype(np.float32) + np.float32(0.5)) / np.float32(2)
    a *= ((theta < (178.0 * np.pi / 180)).astype(np.float32) + np.float32(0.5)) / np.float32(2)
    a *= np.asarray(bands["valid"]).astype(np.float32)
    return a.astype(np.float32)
