"""FP64 oracle of the band-detection path -- TEST INFRASTRUCTURE ONLY.

Same algorithm as the reference CPU path (``pyebsdindex/radon_fast.py:242-310``,
``pyebsdindex/band_detect.py:562-809``) but with no float32 rounding anywhere:

* Radon: ``sum(x - b) / count`` in float64 (the reference rounds every ``x - b`` to
  float32 first, radon_fast.py:258-259),
* padding: written as an index map (the reference does in-place slice copies,
  band_detect.py:573-580); the two are checked against each other in the tests,
* convolution: direct float64 sum over the taps, zero outside the padded array
  (the reference uses a float32 FFT, band_detect.py:605-606),
* local max / labelling: float64 values; every selected peak also reports the gap to
  the next-ranked candidate so that tests can classify near-ties.
"""
import numpy as np

RADEG = 180.0 / np.pi


def radon64(images, indexPlan, background=None):
    """(N,H,W) -> (N, nRho, nTheta) float64; -1 where a line has no pixel."""
    N = images.shape[0]
    nPx = images.shape[1] * images.shape[2]
    x = images.reshape(N, nPx).astype(np.float64)
    if background is not None:
        x = x - np.asarray(background, dtype=np.float64).reshape(1, nPx)
    plan = np.asarray(indexPlan, dtype=np.int64)
    valid = plan < nPx
    cnt = valid.sum(axis=2)
    idx = np.where(valid, plan, 0)
    out = np.empty((N,) + plan.shape[:2], dtype=np.float64)
    for q in range(N):
        s = np.where(valid, x[q][idx], 0.0).sum(axis=2)
        out[q] = np.where(cnt > 0, s / np.maximum(cnt, 1), -1.0)
    return out


def pad_index_map(nR, nT, p0, p1):
    """Row/col source index of every padded cell (band_detect.py:573-580 as a map).

    Returns (src_r, src_c) of shape (nR+2p0, nT+2p1) into the unpadded Radon.
    """
    nRp, nTp = nR + 2 * p0, nT + 2 * p1
    r = np.arange(nRp)
    i = np.clip(r - p0, 0, nR - 1)
    src_r = np.empty((nRp, nTp), dtype=np.int64)
    src_c = np.empty((nRp, nTp), dtype=np.int64)
    for c in range(nTp):
        if c < p1:
            src_r[:, c] = nR - 1 - i
            src_c[:, c] = nT - p1 + c
        elif c >= p1 + nT:
            src_r[:, c] = nR - 1 - i
            src_c[:, c] = c - (p1 + nT)
        else:
            src_r[:, c] = i
            src_c[:, c] = c - p1
    return src_r, src_c


def pad64(radon, p0, p1):
    """(N, nR, nT) -> (N, nRp, nTp) padded copy."""
    N, nR, nT = radon.shape
    sr, sc = pad_index_map(nR, nT, p0, p1)
    return radon[:, sr, sc]


def conv64(padded, kernel):
    """Direct 'same' convolution, zero outside the array.  padded: (N, nRp, nTp)."""
    K = np.asarray(kernel, dtype=np.float64).reshape(kernel.shape[-2], kernel.shape[-1])
    kr, kt = K.shape
    cr, ct = (kr - 1) // 2, (kt - 1) // 2
    N, nRp, nTp = padded.shape
    Z = np.zeros((N, nRp + kr - 1, nTp + kt - 1), dtype=np.float64)
    Z[:, cr:cr + nRp, ct:ct + nTp] = padded
    out = np.zeros((N, nRp, nTp), dtype=np.float64)
    # C[r,c] = sum_{a,b} K[a,b] * P[r + cr - a, c + ct - b]
    for a in range(kr):
        for b in range(kt):
            out += K[a, b] * Z[:, kr - 1 - a:kr - 1 - a + nRp, kt - 1 - b:kt - 1 - b + nTp]
    return out


def stats64(conv, p0, p1):
    """band_detect.py:627-630: rows [p0, nTp-p0) (clipped to nRp), cols [p1, nTp-p1)."""
    N, nRp, nTp = conv.shape
    sl = conv[:, p0:nTp - p0, p1:nTp - p1]
    return sl.min(axis=(1, 2)), sl.mean(axis=(1, 2)), sl.max(axis=(1, 2))


def local_max64(conv, peakPad, rdnmask, p0, p1):
    """band_detect.py:639-664 with scipy's maximum_filter semantics (clamped window)."""
    from scipy.ndimage import maximum_filter
    N, nRp, nTp = conv.shape
    D = maximum_filter(conv, size=(1, int(peakPad[0]), int(peakPad[1])), mode='nearest')
    L = D == conv
    m = np.zeros((nRp, nTp), dtype=bool)
    m[p0:nRp - p0, p1:nTp - p1] = rdnmask
    return L & m[None]


def label64(conv, lmax, nB, p0, p1):
    """band_detect.py:691-809 in float64; returns a dict of (N, nB[,2]) arrays plus
    'gap' = value difference to the next-ranked candidate (inf if none)."""
    N, nRp, nTp = conv.shape
    res = {
        'max': np.full((N, nB), -2.0e6), 'avemax': np.full((N, nB), -2.0e6),
        'maxloc': np.zeros((N, nB, 2)), 'aveloc': np.zeros((N, nB, 2)),
        'valid': np.zeros((N, nB), dtype=np.int8), 'width': np.zeros((N, nB)),
        'gap': np.full((N, nB), np.inf), 'ncand': np.zeros(N, dtype=np.int64),
    }
    for q in range(N):
        C = conv[q]
        rr, cc = np.nonzero(lmax[q])
        vals = C[rr, cc]
        order = np.lexsort((rr * nTp + cc, vals))  # ascending value, ties by position
        res['ncand'][q] = len(order)
        nBq = min(nB, len(order))
        for i in range(nBq):
            k = order[-1 - i]
            if i + 1 < len(order):
                res['gap'][q, i] = vals[k] - vals[order[-2 - i]]
            r, c = int(rr[k]), int(cc[k])
            if not C[r, c] > 0.0:
                continue
            res['maxloc'][q, i] = (r, c)
            res['max'][q, i] = C[r, c]
            col = C[r - 1:r + 2, c]
            if col.min() > 0.0:
                a = (np.log(C[r + 1, c]) - np.log(C[r - 1, c])) * 0.5 - np.log(C[r, c])
                if a < -1.e-8:
                    res['width'][q, i] = 2.0 * np.sqrt(np.log(2.0) / -a)
            nn = C[r - 1:r + 2, c - 1:c + 2].copy()
            s = nn.sum()
            s = s if s > 1e-12 else 1e-12
            nn /= s
            res['avemax'][q, i] = s / 9.0
            dx = 0.5 * (nn[1, 2] - nn[1, 0])
            dy = 0.5 * (nn[2, 1] - nn[0, 1])
            dxx = nn[1, 2] + nn[1, 0] - 2 * nn[1, 1]
            dyy = nn[2, 1] + nn[0, 1] - 2 * nn[1, 1]
            dxy = 0.25 * (nn[2, 2] - nn[0, 2] - nn[2, 0] + nn[0, 0])
            det = dxx * dyy - dxy * dxy
            det = det if abs(det) > 1e-12 else 1e-12
            dc = (dyy * dx - dxy * dy) / det
            rc = (dxx * dy - dxy * dx) / det
            if abs(dc) > 0.875 or abs(rc) > 0.875:
                det = dxx * dyy
                det = det if abs(det) > 1e-12 else 1e-12
                dc = (dyy * dx) / det
                rc = (dxx * dy) / det
                if abs(dc) > 0.875 or abs(rc) > 0.875:
                    dc = rc = 0.0
            res['aveloc'][q, i] = (r - rc, c - dc)
            res['valid'][q, i] = 1
    res['maxloc'] -= np.array([p0, p1])
    res['aveloc'] -= np.array([p0, p1])
    return res


def find_bands64(images, indexPlan, kernel, rdnmask, padding, peakPad, nB, theta_deg, rho_grid,
                 background=None):
    """Whole path in float64; returns (label dict, stats tuple, conv, radon)."""
    p0, p1 = int(padding[0]), int(padding[1])
    R = radon64(images, indexPlan, background)
    P = pad64(R, p0, p1)
    C = conv64(P, kernel)
    st = stats64(C, p0, p1)
    L = local_max64(C, peakPad, rdnmask, p0, p1)
    res = label64(C, L, nB, p0, p1)
    mn, ave = st[0][:, None], st[1][:, None]
    res['normmax'] = (res['max'] - mn) / np.clip(ave - mn, 1e-7, None)
    nT, nR = len(theta_deg), len(rho_grid)
    res['theta'] = np.pi - np.interp(res['aveloc'][:, :, 1], np.arange(nT), theta_deg) / RADEG
    res['rho'] = -1.0 * np.interp(res['aveloc'][:, :, 0], np.arange(nR), rho_grid)
    return res, st, C, R
