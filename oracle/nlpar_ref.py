"""NLPAR oracle -- TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``sigma`` / ``nlpar``: the reference's two Numba kernels restated in C
  (``csrc/ref_nlpar.c``: pyebsdindex/nlpar_cpu.py:753-819 and :821-935), float32 where
  the reference is float32; also the timed CPU baseline.
* ``sigma64`` / ``nlpar64``: the same definitions in float64 numpy, written from the
  NLPAR statistic (Brewick, Wright, Rowenhorst, Ultramicroscopy 200 (2019) 50-61), for
  small scans: the normalised distance (d^2 - n (s_a^2 + s_b^2)) / ((s_a^2 + s_b^2)
  sqrt(2 n)), weights exp(-max(d - dthresh, 0) / lambda^2) with weight 1 on the centre.
* ``makeautomask``: nlpar_cpu.py:737-751.
* ``calcnlpar_options``: calcnlpar's data path with its options (nlpar_cpu.py:470-526):
  stem_scale (file minimum, sqrt, square back), backsub (``background`` of the scan's mean,
  ``backsub_fit`` = NLPAR.backsub :679-735), rescale, then the clip to the output type.

How they are pinned: tests/test_oracle_nlpar.py (closed forms: constant-offset
neighbours give sigma = |c| / sqrt(2); lambda -> infinity gives the shifted-window mean,
lambda -> 0 the pattern itself; saturated pixels drop out of the counts; C against
float64 within float32 rounding).  The options (tests/test_oracle_nlpar_options.py):
lambda -> 0 with stem_scale returns the patterns within one count; rescale maps every
pattern onto [0, top]; numpy's mean over the patterns is the sequential float32 sum over
them; the background fit recovers an exact Gaussian-on-a-plane surface.
"""
import ctypes

import numpy as np

from . import ref_cpu


def _lib():
    L = ref_cpu.lib()
    if not hasattr(L, "_nlpar_set"):
        P, I, F = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float
        L.ora_nlpar_sigma.argtypes = [P, I, I, I, P, I, I, ctypes.c_int, P, P, P]
        L.ora_nlpar_sigma.restype = None
        L.ora_nlpar.argtypes = [P, F, I, F, P, I, I, I, P, I, ctypes.c_int, F, I, I, I, I, P, P]
        L.ora_nlpar.restype = None
        L._nlpar_set = True
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def makeautomask(h, w):
    """Circular mask of radius 0.98 * min(h, w) / 2 about the pattern centre
    (nlpar_cpu.py:737-751)."""
    r = (min(h, w) * 0.98 * 0.5)
    x = np.arange(w, dtype=np.float32)
    x = np.minimum(x, (w - x)).reshape(1, w)
    y = np.arange(h, dtype=np.float32)
    y = np.minimum(y, (h - y)).reshape(h, 1)
    m = (np.sqrt(y ** 2 + x ** 2) < r).astype(np.uint8)
    return np.roll(m, (int(h / 2), int(w / 2)), axis=(0, 1))


def sigma(data, nrows, ncols, indices, nn=1, saturation_protect=True):
    """(sigma (nrows, ncols), dout, nout (nrows, ncols, (2nn+1)^2)) from the C restatement.
    ``data``: (nrows*ncols, P) float32."""
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(nrows * ncols, -1)
    P = d.shape[1]
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    win = (2 * nn + 1) ** 2
    s = np.zeros(nrows * ncols, np.float32)
    dout = np.zeros(nrows * ncols * win, np.float32)
    nout = np.zeros(nrows * ncols * win, np.float32)
    _lib().ora_nlpar_sigma(_p(d), nrows, ncols, P, _p(idx), idx.size, nn, int(bool(saturation_protect)),
                           _p(s), _p(dout), _p(nout))
    return s.reshape(nrows, ncols), dout.reshape(nrows, ncols, win), nout.reshape(nrows, ncols, win)


def nlpar(data, nrows, ncols, indices, sig, lam=0.7, sr=3, dthresh=0.0, saturation_protect=True,
          diff_offset=0.0, calclim=None, return_weights=False):
    """Averaged patterns (nrows*ncols, P) float32 from the C restatement (points outside
    ``calclim`` = (cstart, rstart, ncols, nrows) stay 0)."""
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(nrows * ncols, -1)
    P = d.shape[1]
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    s = np.ascontiguousarray(sig, dtype=np.float32).reshape(-1)
    cl = (0, 0, ncols, nrows) if calclim is None else tuple(int(v) for v in calclim)
    out = np.zeros((nrows * ncols, P), np.float32)
    w = np.zeros(nrows * ncols * (2 * sr + 1) ** 2, np.float32) if return_weights else None
    _lib().ora_nlpar(_p(d), float(lam), sr, float(dthresh), _p(s), nrows, ncols, P, _p(idx), idx.size,
                     int(bool(saturation_protect)), float(diff_offset), cl[0], cl[1], cl[2], cl[3], _p(out),
                     None if w is None else _p(w))
    if return_weights:
        return out, w.reshape(nrows * ncols, -1)
    return out


def _window(j, i, nrows, ncols, sr):
    """Shifted (2sr+1)^2 window of point (j, i), row-major (nlpar_cpu.py:872-876)."""
    y0 = min(max(j - sr, 0), nrows - (2 * sr + 1))
    x0 = min(max(i - sr, 0), ncols - (2 * sr + 1))
    return [(y, x) for y in range(y0, y0 + 2 * sr + 1) for x in range(x0, x0 + 2 * sr + 1)]


def _pair(a, b, sat):
    ok = (a < sat) & (b < sat)
    return float(np.sum((a[ok] - b[ok]) ** 2)), int(ok.sum())


def sigma64(data, nrows, ncols, indices, nn=1, saturation_protect=True):
    """float64 sigma: sqrt of the smallest d^2 / 2n over the +-nn neighbours (clipped)
    whose d^2 >= 1e-3."""
    d = np.asarray(data, np.float64).reshape(nrows * ncols, -1)[:, indices]
    mx = d.max() * 0.9961 if saturation_protect else d.max() + 1.0
    out = np.zeros((nrows, ncols))
    for j in range(nrows):
        for i in range(ncols):
            best = 1e24
            for y in range(max(j - nn, 0), min(j + nn, nrows - 1) + 1):
                for x in range(max(i - nn, 0), min(i + nn, ncols - 1) + 1):
                    if (y, x) == (j, i):
                        continue
                    d2, n2 = _pair(d[j * ncols + i], d[y * ncols + x], mx)
                    if d2 >= 1e-3:
                        best = min(best, d2 / (2.0 * n2))
            out[j, i] = np.sqrt(best)
    return out


def nlpar64(data, nrows, ncols, indices, sig, lam=0.7, sr=3, dthresh=0.0, saturation_protect=True):
    """float64 NLPAR of every point (whole scan, no chunks)."""
    full = np.asarray(data, np.float64).reshape(nrows * ncols, -1)
    d = full[:, indices]
    mx = d.max() * 0.999 if saturation_protect else d.max() + 1.0
    s2 = np.asarray(sig, np.float64).reshape(-1) ** 2
    out = np.zeros_like(full)
    for j in range(nrows):
        for i in range(ncols):
            p = j * ncols + i
            ws, qs = [], []
            for (y, x) in _window(j, i, nrows, ncols, sr):
                q = y * ncols + x
                if q == p:
                    w = 1.0
                else:
                    d2, n2 = _pair(d[p], d[q], mx)
                    s12 = s2[p] + s2[q]
                    dn = s12 * np.sqrt(2.0 * n2)
                    dist = (d2 - n2 * s12) / dn if dn > 1e-8 else 1e6 * n2
                    w = np.exp(-max(dist - dthresh, 0.0) / lam ** 2)
                ws.append(w)
                qs.append(q)
            ws = np.array(ws) / np.sum(ws)
            out[p] = ws @ full[qs]
    return out


def backsub_fit(mean, mask):
    """NLPAR.backsub's fit (nlpar_cpu.py:679-735) to a mean pattern: a 2-D Gaussian on a
    plane by scipy's curve_fit at the mask's pixels, from the reference's starting point,
    minus its mean; the zero-mean mean pattern when the fit does not converge."""
    import scipy.optimize as opt
    back = np.asarray(mean)
    ny, nx = back.shape
    yy, xx = np.mgrid[0:ny, 0:nx]
    x, y = xx.ravel().astype(float), yy.ravel().astype(float)

    def g(M, a, x0, y0, sx, sy, c, d, e):
        return a * np.exp(-((M[0] - x0) ** 2) / (2.0 * sx ** 2) - ((M[1] - y0) ** 2) / (2.0 * sy ** 2)) \
            + c + d * M[0] + e * M[1]

    sel = np.flatnonzero(np.asarray(mask).ravel())
    zz = back.ravel()[sel]
    iy, ix = np.unravel_index(np.argmax(back), back.shape)
    start = [zz.max() - zz.min(), ix, iy, nx / 2.355, ny / 2.355, zz.min(), 0, 0]
    try:
        par, _ = opt.curve_fit(g, np.vstack((x[sel], y[sel])), zz, start)
        surface = g(np.vstack((x, y)), *par).reshape(ny, nx)
    except RuntimeError:
        surface = back.copy()
    return surface - np.mean(surface)


def calcnlpar_options(pats, nrows, ncols, indices, lam=0.7, sr=3, stem_scale=False, backsub=False,
                      rescale_top=None, out_dtype=np.uint8, saturation_protect=True, mask=None):
    """The averaged patterns calcnlpar writes for a scan that fits one chunk
    (nlpar_cpu.py:470-526) with its options, from the C restatement of the kernels: the
    file's values as float32; stem_scale: sqrt(v - min); backsub: minus ``backsub_fit`` of
    the mean pattern (numpy's float32 mean over the patterns); sigma and the averages on
    that data; stem_scale: squared plus the minimum; rescale: each pattern shifted to 0 and
    scaled by float32(top) / max; clipped to the output type and truncated."""
    d = np.asarray(pats).astype(np.float32).reshape(nrows * ncols, -1)
    H, W = np.asarray(pats).shape[-2:]
    dmin = None
    if stem_scale:
        dmin = d.min()
        d = np.sqrt(d - dmin).astype(np.float32)
    if backsub:
        mean = np.mean(d.reshape(-1, H, W), axis=0)
        fit = backsub_fit(mean, mask)
        d = d.reshape(-1, H, W)
        d -= fit
        d = d.reshape(nrows * ncols, -1)
    sg = sigma(d, nrows, ncols, indices, nn=1, saturation_protect=saturation_protect)[0]
    out = nlpar(d, nrows, ncols, indices, sg, lam=lam, sr=sr, saturation_protect=saturation_protect)
    out = np.asarray(out, dtype=np.float32).reshape(nrows * ncols, H, W)
    if stem_scale:
        out = out ** 2 + dmin
    if rescale_top is not None:
        for i in range(out.shape[0]):
            t = out[i]
            t -= t.min()
            t *= np.float32(rescale_top) / t.max()
    top = np.iinfo(out_dtype).max
    return out.clip(0, top).astype(out_dtype)
