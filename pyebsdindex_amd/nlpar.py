"""NLPAR -- non-local pattern averaging -- on the GPU, behind the reference's ``NLPAR`` API.

Mirrors ``pyebsdindex.nlpar_cpu.NLPAR`` (pyebsdindex/nlpar_cpu.py:47-1072): the same
constructor keywords and attributes, ``calcsigma``, ``opt_lambda``, ``calcnlpar`` and
``auto_nlpar`` over a pattern file, ``setfile`` / ``setoutfile`` / ``getinfileobj`` /
``getoutfileobj`` and ``makeautomask``.  The two Numba kernels run as HIP kernels
(``libebsd_nlpar.so``, include/ebsd_nlpar.h); the lambda optimiser stays on the host with
the same SciPy call.  ``sigma_device`` and ``nlpar_device`` run the kernels on patterns
already on the device.

Scans are processed in bands of rows (the whole scan when it fits the memory budget).
The noise estimate of every point is computed first, over the whole scan (bands overlap
by two rows), and the averaging then uses those values, with bands overlapping by the
search radius so that each computed point sees its full window.  This gives the
reference's result for a scan that fits one of its chunks.  When the reference splits a
scan into several chunks, the points next to a chunk boundary use the noise estimate of
the chunk processed so far (nlpar_cpu.py:509-516); that order-dependent value is not
reproduced.  The saturation threshold is taken over the whole scan.

The options of calcnlpar run on the device too (nlpar_cpu.py:470-526):
``stem_scale`` takes sqrt(v - file minimum) before the averaging and squares back after;
``backsub`` subtracts a Gaussian-on-a-plane fitted (on the host, with SciPy as the
reference) to the mean pattern; ``rescale`` stretches each averaged pattern to the output
type's range.  The background is fitted once, to the mean of the whole scan: the reference
fits each of its chunks' means, which is the same for a scan that fits one chunk.
``diff_offset`` is added once to each pattern's own weight; the reference keeps it in an
array that all of its parallel workers share and that grows along a column
(nlpar_cpu.py:890-892), which is not reproduced (the default, 0, is unaffected).
"""
import ctypes
from pathlib import Path

import numpy as np

from . import _hip
from . import ebsd_pattern

def _torch():
    import torch
    return torch


def _code(t):
    torch = _torch()
    m = {torch.uint8: 0, torch.uint16: 1, torch.float32: 2}
    if t.dtype not in m:
        raise TypeError(f"patterns must be uint8, uint16 or float32, not {t.dtype}")
    return m[t.dtype]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _indices_dev(indices, P, device):
    """Mask pixel indices as device int32, checked against the pattern size."""
    torch = _torch()
    idx = np.asarray(indices, dtype=np.int64).reshape(-1)
    if idx.size and (idx.min() < 0 or idx.max() >= P):
        raise ValueError("mask pixel indices must lie in [0, H*W)")
    return torch.as_tensor(idx.astype(np.int32), device=device)


def _workspace(x, pairs=None):
    """(buffer, bytes) of the workspace for the chunk ``x`` (n, P): the integer-pattern part
    (none for float32 patterns), and with ``pairs`` = (calculation points, search radius)
    room for ebsd_nlpar's pair table after it."""
    torch = _torch()
    L = _hip.load_nlpar()
    base = int(L.ebsd_nlpar_workspace_bytes(x.shape[0], x.shape[1]))
    if pairs is not None:
        nbytes = -(-base // 256) * 256 + int(L.ebsd_nlpar_pair_bytes(int(pairs[0]), int(pairs[1])))
    elif x.dtype not in (torch.uint8, torch.uint16):
        return None, 0
    else:
        nbytes = base
    buf = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    return buf, nbytes


def _stream(stream, device=None):
    """The launch stream as a handle: the current stream, or ``stream`` after it has waited
    for the current stream's queued work (see _hip.launch_stream)."""
    s = _hip.launch_stream(stream, device)
    return ctypes.c_void_p(s.cuda_stream)


def _sat_sigma(mx, saturation_protect):
    """sigma_numba's threshold (nlpar_cpu.py:763-767): float64 max * 0.9961 (or + 1), as
    the smallest float32 not below it, so float32 comparisons give the float64 result."""
    t = float(mx) * 0.9961 if saturation_protect else float(mx) + 1.0
    f = np.float32(t)
    if float(f) < t:
        f = np.nextafter(f, np.float32(np.inf))
    return float(f)


def _sat_nlpar(mx, saturation_protect):
    """nlpar_nb's threshold (nlpar_cpu.py:866-870): float32 max * float32(0.999) (or + 1)."""
    m = np.float32(mx)
    return float(m * np.float32(0.999) if saturation_protect else m + np.float32(1.0))


def makeautomask(h, w):
    """Circular mask of radius 0.98 * min(h, w) / 2 about the pattern centre
    (nlpar_cpu.py:737-751)."""
    r = (min(h, w) * 0.98 * 0.5)
    x = np.arange(w, dtype=np.float32)
    x = np.minimum(x, (w - x)).reshape(1, w)
    y = np.arange(h, dtype=np.float32)
    y = np.minimum(y, (h - y)).reshape(h, 1)
    mask = (np.sqrt(y ** 2 + x ** 2) < r).astype(np.uint8)
    return np.roll(mask, (int(h / 2), int(w / 2)), axis=(0, 1))


def pattern_max(pats, stream=None):
    """Largest pixel value of device patterns, as float32 (a device reduction)."""
    torch = _torch()
    L = _hip.load_nlpar()
    pats = pats.contiguous()
    out = torch.empty(1, dtype=torch.float32, device=pats.device)
    n = pats.shape[0] if pats.dim() > 2 else 1
    P = pats.numel() // max(n, 1)
    _hip.check_nlpar(L.ebsd_nlpar_max(_ptr(pats), _code(pats), n, P, _ptr(out), _stream(stream, pats.device)))
    if stream is not None:
        stream.synchronize()   # the reduction ran on ``stream``: done before the read
    return float(out.item())


def pattern_min(pats, stream=None):
    """Smallest pixel value of device patterns, as float32 (a device reduction)."""
    torch = _torch()
    L = _hip.load_nlpar()
    pats = pats.contiguous()
    out = torch.empty(1, dtype=torch.float32, device=pats.device)
    n = pats.shape[0] if pats.dim() > 2 else 1
    P = pats.numel() // max(n, 1)
    _hip.check_nlpar(L.ebsd_nlpar_min(_ptr(pats), _code(pats), n, P, _ptr(out), _stream(stream, pats.device)))
    if stream is not None:
        stream.synchronize()
    return float(out.item())


def prepare_device(pats, stem_min=None, background=None, stream=None):
    """The float32 patterns calcnlpar averages (nlpar_cpu.py:470-482): ``stem_min`` (a
    float32 value, or None): sqrt(v - stem_min); ``background`` ((H, W) float64 device
    tensor, or None): minus it.  Returns a new float32 device tensor of the same shape."""
    torch = _torch()
    L = _hip.load_nlpar()
    x = pats.contiguous()
    n = x.shape[0] if x.dim() > 2 else 1
    P = x.numel() // max(n, 1)
    out = torch.empty(x.shape, dtype=torch.float32, device=x.device)
    bg = None
    if background is not None:
        bg = background.to(torch.float64).contiguous().reshape(-1)
        if bg.numel() != P:
            raise ValueError("the background must have one value per pattern pixel")
    _hip.check_nlpar(L.ebsd_nlpar_prepare(_ptr(x), _code(x), n, P, int(stem_min is not None),
                                          float(np.float32(0.0 if stem_min is None else stem_min)),
                                          None if bg is None else _ptr(bg), _ptr(out), _stream(stream, x.device)))
    if stream is not None:
        _hip.keep_for_stream(stream, x, bg, out)
    return out


def finish_device(x, stem_min=None, rescale_top=None, out_dtype=None, stream=None):
    """Averaged float32 device patterns (n, H, W) to the output values (nlpar_cpu.py:511-526):
    squared plus ``stem_min`` when given, stretched to [0, ``rescale_top``] per pattern when
    given, then ``out_dtype`` (torch.float32, or uint8/uint16 clipped and truncated)."""
    torch = _torch()
    L = _hip.load_nlpar()
    x = x.contiguous()
    n = x.shape[0] if x.dim() > 2 else 1
    P = x.numel() // max(n, 1)
    odt = torch.float32 if out_dtype is None else out_dtype
    out = torch.empty(x.shape, dtype=odt, device=x.device)
    _hip.check_nlpar(L.ebsd_nlpar_finish(_ptr(x), n, P, int(stem_min is not None),
                                         float(np.float32(0.0 if stem_min is None else stem_min)),
                                         float(rescale_top or 0.0), _ptr(out), _code(out), _stream(stream, x.device)))
    if stream is not None:
        _hip.keep_for_stream(stream, x, out)
    return out


def fit_background(mean, mask):
    """NLPAR.backsub's background (nlpar_cpu.py:679-735): a 2-D Gaussian on a plane fitted by
    least squares to the mean pattern at the mask's pixels, started from (max - min, the
    maximum's position, size / 2.355, min, flat), then shifted to zero mean; the mean pattern
    itself (zero-mean) if the fit does not converge.  Returns float64 (H, W)."""
    import scipy.optimize as opt
    back = np.asarray(mean)
    ny, nx = back.shape
    x = np.broadcast_to(np.arange(nx, dtype=float).reshape(1, nx), (ny, nx)).ravel()
    y = np.broadcast_to(np.arange(ny, dtype=float).reshape(ny, 1), (ny, nx)).ravel()

    def surf(x, y, a, x0, y0, sx, sy, c, d, e):
        return a * np.exp(-((x - x0) ** 2) / (2.0 * sx ** 2) - ((y - y0) ** 2) / (2.0 * sy ** 2)) + c + d * x + e * y

    wh = np.nonzero(np.asarray(mask).ravel())[0]
    z = back.ravel()[wh]
    top = np.unravel_index(back.argmax(), back.shape)
    p0 = [(z.max() - z.min()), top[1], top[0], nx / 2.355, ny / 2.355, z.min(), 0, 0]
    try:
        popt, _ = opt.curve_fit(lambda M, *a: surf(M[0], M[1], *a), np.vstack((x[wh], y[wh])), z, p0)
        fit = surf(x, y, *popt).reshape(ny, nx)
    except RuntimeError:
        print('Warning: no convergence on back subtract ... using mean of the patterns.')
        fit = back.copy()
    fit -= np.mean(fit)
    return fit


def sigma_device(pats, nrows, ncols, indices, nn=1, saturation_protect=True, return_d2=False, mxval=None,
                 stream=None, fast=True):
    """Noise estimate of every point (nlpar_cpu.py:753-819) for device patterns
    (nrows*ncols, H, W) or (nrows, ncols, H, W).  Returns sigma (nrows, ncols) and, with
    ``return_d2``, the normalised distances and counts (nrows, ncols, (2nn+1)^2), all device
    float32.  ``mxval``: the saturation reference value (default: the patterns' maximum).
    ``fast``: uint8/uint16 distances as exact integer sums over packed words (a workspace
    of about one pattern plus 265 bytes per pattern); False: pixel by pixel."""
    torch = _torch()
    L = _hip.load_nlpar()
    x = pats.contiguous().reshape(nrows * ncols, -1)
    P = x.shape[1]
    idx = _indices_dev(indices, P, x.device)
    mx = pattern_max(x, stream) if mxval is None else float(mxval)
    sig = torch.empty((nrows, ncols), dtype=torch.float32, device=x.device)
    win = (2 * nn + 1) ** 2
    dout = nout = None
    if return_d2:
        dout = torch.empty((nrows, ncols, win), dtype=torch.float32, device=x.device)
        nout = torch.empty_like(dout)
    work, wbytes = _workspace(x) if fast else (None, 0)
    _hip.check_nlpar(L.ebsd_nlpar_sigma(_ptr(x), _code(x), nrows, ncols, P, _ptr(idx), idx.numel(), nn,
                                        _sat_sigma(mx, saturation_protect), _ptr(sig),
                                        None if dout is None else _ptr(dout), None if nout is None else _ptr(nout),
                                        None if work is None else _ptr(work), wbytes, _stream(stream, x.device)))
    if stream is not None:
        _hip.keep_for_stream(stream, x, idx, sig, dout, nout, work)
    return (sig, dout, nout) if return_d2 else sig


def nlpar_device(pats, nrows, ncols, indices, sigma, lam=0.7, sr=3, dthresh=0.0, saturation_protect=True,
                 diff_offset=0.0, calclim=None, out_dtype=None, return_weights=False, mxval=None, stream=None,
                 fast=True):
    """Averaged patterns (nlpar_cpu.py:821-935) of the points in ``calclim`` = (cstart,
    rstart, ncolcalc, nrowcalc) (default: all) of device patterns (nrows*ncols, H, W) or
    (nrows, ncols, H, W), as (nrowcalc*ncolcalc, H, W) of ``out_dtype`` (a torch dtype:
    float32 (default), or uint8 / uint16 clipped and truncated as the reference writes
    them).  ``sigma``: (nrows, ncols) device float32.  ``fast``: as for ``sigma_device``."""
    torch = _torch()
    L = _hip.load_nlpar()
    x = pats.contiguous().reshape(nrows * ncols, -1)
    P = x.shape[1]
    H, W = (pats.shape[-2], pats.shape[-1]) if pats.dim() >= 3 else (1, P)
    idx = _indices_dev(indices, P, x.device)
    sig = torch.as_tensor(sigma, dtype=torch.float32, device=x.device).contiguous().reshape(-1)
    if sig.numel() != nrows * ncols:
        raise ValueError("sigma must have one value per scan point")
    c0, r0, ncc, nrc = (0, 0, ncols, nrows) if calclim is None else (int(v) for v in calclim)
    odt = torch.float32 if out_dtype is None else out_dtype
    out = torch.empty((nrc * ncc, H, W), dtype=odt, device=x.device)
    wts = torch.empty((nrc * ncc, (2 * sr + 1) ** 2), dtype=torch.float32, device=x.device) \
        if return_weights else None
    mx = pattern_max(x, stream) if mxval is None else float(mxval)
    work, wbytes = _workspace(x, pairs=(nrc * ncc, sr)) if fast else (None, 0)
    _hip.check_nlpar(L.ebsd_nlpar(_ptr(x), _code(x), nrows, ncols, P, _ptr(idx), idx.numel(), _ptr(sig), int(sr),
                                  float(lam), float(dthresh), float(diff_offset), _sat_nlpar(mx, saturation_protect),
                                  r0, c0, nrc, ncc, _ptr(out), _code(out), None if wts is None else _ptr(wts),
                                  None if work is None else _ptr(work), wbytes, _stream(stream, x.device)))
    if stream is not None:
        _hip.keep_for_stream(stream, x, idx, sig, out, wts, work)
    return (out, wts) if return_weights else out


def _scalar(v):
    return v if np.isscalar(v) else v[0]


class NLPAR:
    """GPU NLPAR with the reference's interface (nlpar_cpu.py:47-79)."""

    def __init__(self, filename=None, lam=0.7, searchradius=3, dthresh=0.0, diff_offset=0.0,
                 saturation_protect=True, stem_scale=False, automask=True, nrows=None, ncols=None,
                 gpu_id=None, band_bytes=None, **kwargs):
        self.lam = lam
        self.searchradius = searchradius
        self.dthresh = dthresh
        self.diff_offset = diff_offset
        self.saturation_protect = saturation_protect
        self.stem_scale = stem_scale
        self.automask = automask
        self.filepath = None
        self.hdfdatapath = None
        self.filepathout = None
        self.hdfdatapathout = None
        self.setfile(filename)
        self.mask = None
        self.sigma = None
        self.sigmann = 1
        self.nrows = nrows
        self.ncols = ncols
        self.gpu_id = gpu_id
        #: device bytes of patterns per band (default: a quarter of the free device memory)
        self.band_bytes = band_bytes

    # -- files (nlpar_cpu.py:542-676) ------------------------------------------------------
    def setfile(self, filepath=None):
        self.filepath = None
        self.hdfdatapath = None
        if filepath is not None:
            self.filepath = Path(np.atleast_1d(filepath)[0])

    def setoutfile(self, patternfile, filepath=None):
        """The output file: ``filepath``, or ``<stem>_NLPAR_l{lam:.2f}sr{sr}<suffix>`` beside
        the input; a header copy of the input with room for the patterns."""
        self.filepathout = None
        self.hdfdatapathout = None
        if filepath is not None:
            out = Path(np.atleast_1d(filepath)[0]).expanduser().resolve()
            if out.exists() and out.samefile(patternfile.filepath):
                raise ValueError('Error: File input and output are exactly the same.')
        else:
            p = Path(patternfile.filepath)
            appnd = "_NLPAR_l{:1.2f}".format(float(_scalar(self.lam))) + "sr{:d}".format(int(_scalar(self.searchradius)))
            out = p.parent / (p.stem + appnd + p.suffix)
        if getattr(patternfile, "filetype", "UP") != "UP":
            # .ebsp / .pat: a copy of the input whose patterns are then overwritten
            # (nlpar_cpu.py:590,603-613; TF copies keep only the header and background)
            patternfile.copy_file(str(out), empty_data=True)
            self.filepathout = out
            return
        w = ebsd_pattern.UPFile(out)
        for k in ("version", "patternW", "patternH", "filePos", "extraPatterns", "nCols", "nRows", "hexflag",
                  "xStep", "yStep", "filedatatype"):
            setattr(w, k, getattr(patternfile, k))
        if w.version == 1:
            w.version = 3
            w.filePos = 42
            w.nCols, w.nRows = patternfile.nCols, patternfile.nRows
        w.write_header()
        with open(out, 'r+b') as f:
            f.truncate(int(w.filePos) + int(patternfile.nPatterns) * w.pattern_bytes)
        self.filepathout = out

    def getinfileobj(self):
        if self.filepath is None:
            return None
        f = ebsd_pattern.get_pattern_file_obj(str(self.filepath))
        if f.nRows is not None:
            if self.nrows is None:
                self.nrows = f.nRows
            else:
                f.nRows = self.nrows
        if f.nCols is not None:
            if self.ncols is None:
                self.ncols = f.nCols
            else:
                f.nCols = self.ncols
        return f

    def getoutfileobj(self):
        if self.filepathout is None:
            return None
        f = ebsd_pattern.get_pattern_file_obj(str(self.filepathout))
        if self.nrows is not None:
            f.nRows = self.nrows
        else:
            self.nrows = f.nRows
        if self.ncols is not None:
            f.nCols = self.ncols
        else:
            self.ncols = f.nCols
        return f

    @staticmethod
    def makeautomask(h, w):
        return makeautomask(h, w)

    # -- helpers ---------------------------------------------------------------------------
    def _device(self):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _hip.HipLibraryError("NLPAR needs a HIP device")
        return torch.device("cuda", torch.cuda.current_device() if self.gpu_id is None else int(self.gpu_id))

    def _indices(self, H, W, automask):
        if automask and self.mask is None:
            self.mask = makeautomask(H, W)
        if self.mask is None:
            self.mask = np.ones((H, W), dtype=np.uint8)
        return np.asarray(self.mask.flatten().nonzero()[0], np.int32)

    def _options(self, f, dev, stem, backsub, bands):
        """The device transform of the patterns for stem_scale and backsub: the file's
        smallest value (stem_scale, nlpar_cpu.py:948-969) and the background fitted to the
        mean of the (stem-scaled) scan (backsub, :679-735).  ``bands``: row ranges that
        cover the scan once."""
        torch = _torch()
        self._rows_cache = None
        self._opt = None
        if not (stem or backsub):
            return
        dmin = None
        if stem:
            m = np.inf
            for r0, r1 in bands:
                m = min(m, pattern_min(self._read_rows(f, r0, r1, dev)))
            dmin = np.float32(m)
        back = None
        if backsub:
            H, W = int(f.patternH), int(f.patternW)
            L = _hip.load_nlpar()
            acc = torch.zeros(H * W, dtype=torch.float32, device=dev)
            for r0, r1 in bands:
                x = prepare_device(self._read_rows(f, r0, r1, dev), stem_min=dmin)
                _hip.check_nlpar(L.ebsd_nlpar_colsum(_ptr(x), x.shape[0], H * W, _ptr(acc), _stream(None, dev)))
            n = int(self.nrows) * int(self.ncols)
            # np.mean divides the float32 sums by the count in float64 and rounds to float32
            mean = (acc.cpu().numpy().astype(np.float64) / n).astype(np.float32).reshape(H, W)
            back = torch.as_tensor(fit_background(mean, self.mask), dtype=torch.float64, device=dev)
        self._opt = (dmin, back)

    def _bands(self, f, halo, float_rows=False):
        """Row bands (r0, r1) covering the scan, sized to the device budget (``float_rows``:
        the patterns are held as float32)."""
        torch = _torch()
        nrows, ncols = int(self.nrows), int(self.ncols)
        row_bytes = ncols * (4 * int(f.patternW) * int(f.patternH) if float_rows else f.pattern_bytes)
        budget = self.band_bytes
        if budget is None:
            free, _ = torch.cuda.mem_get_info(self._device())
            budget = free // 4
        rows = max(1, int(budget // max(row_bytes, 1)) - 2 * halo)
        return [(r, min(nrows, r + rows)) for r in range(0, nrows, rows)]

    def _load_rows(self, f, a, b, dev):
        """Rows [a, b) of the scan as device patterns (file type), through pinned memory.
        The last rows loaded are kept, so a scan that fits one band is read from the file
        once for the maximum, the sigma and the averaging passes."""
        opt = getattr(self, "_opt", None)
        key = (str(f.filepath), a, b, str(dev), None if opt is None else id(opt))
        cached = getattr(self, "_rows_cache", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        self._rows_cache = None
        x = self._read_rows(f, a, b, dev)
        if opt is not None:
            x = prepare_device(x, stem_min=opt[0], background=opt[1])
        self._rows_cache = (key, x)
        return x

    def _read_rows(self, f, a, b, dev):
        torch = _torch()
        ncols = int(self.ncols)
        tdt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.uint16}.get(np.dtype(f.filedatatype))
        if tdt is None:
            raise ValueError(f"NLPAR reads uint8 and uint16 pattern files, not {np.dtype(f.filedatatype)}")
        host = torch.empty(((b - a) * ncols, int(f.patternH), int(f.patternW)), dtype=tdt, pin_memory=True)
        f.read_into(host, a * ncols, (b - a) * ncols)
        return host.to(dev, non_blocking=False)

    def _scan_max(self, f, bands, dev):
        m = -np.inf
        for r0, r1 in bands:
            m = max(m, pattern_max(self._load_rows(f, r0, r1, dev)))
        return m

    # -- sigma (nlpar_cpu.py:186-281) ------------------------------------------------------
    def calcsigma(self, chunksize=0, nn=1, saturation_protect=None, automask=None, stem_scale=None, verbose=0,
                  _keep_options=False, **kwargs):
        """Noise estimate of every scan point; returns (sigma, d2, n2) as numpy arrays
        ((nrows, ncols), and (nrows, ncols, (2nn+1)^2) normalised distances and counts).
        ``chunksize`` > 0 bounds the device bytes of patterns per band, as the reference's
        chunk target bounds its host chunks (nlpar_cpu.py:218-223)."""
        if chunksize and chunksize > 0:
            self.band_bytes = int(chunksize)
        self.sigmann = nn
        if saturation_protect is not None:
            self.saturation_protect = saturation_protect
        if automask is not None:
            self.automask = automask
        if stem_scale is not None:
            self.stem_scale = stem_scale
        stem = _scalar(self.stem_scale) is True
        f = self.getinfileobj()
        dev = self._device()
        nrows, ncols = int(self.nrows), int(self.ncols)
        idx = self._indices(int(f.patternH), int(f.patternW), _scalar(self.automask))
        # bands overlap by 2 nn rows: the neighbours of a kept point then have their whole
        # neighbourhood in the band too, so the sigmas that normalise its distances are exact
        keep = _keep_options and getattr(self, "_opt", None) is not None
        bands = self._bands(f, 2 * nn, float_rows=stem or keep)
        if not keep:
            self._options(f, dev, stem, False, bands)
        mx = self._scan_max(f, bands, dev)
        win = (2 * nn + 1) ** 2
        sigma = np.zeros((nrows, ncols), np.float32)
        d2 = np.zeros((nrows, ncols, win), np.float32)
        n2 = np.zeros((nrows, ncols, win), np.float32)
        for r0, r1 in bands:
            a, b = max(0, r0 - 2 * nn), min(nrows, r1 + 2 * nn)
            x = self._load_rows(f, a, b, dev)
            s, dd, cnt = sigma_device(x, b - a, ncols, idx, nn=nn, saturation_protect=_scalar(self.saturation_protect),
                                      return_d2=True, mxval=mx)
            sigma[r0:r1] = s[r0 - a:r1 - a].cpu().numpy()
            d2[r0:r1] = dd[r0 - a:r1 - a].cpu().numpy()
            n2[r0:r1] = cnt[r0 - a:r1 - a].cpu().numpy()
        if verbose >= 1:
            print(sigma.min(), sigma.max())
        if not keep:
            self._opt = None
            self._rows_cache = None
        return sigma, d2, n2

    # -- lambda (nlpar_cpu.py:94-184) ------------------------------------------------------
    def opt_lambda(self, target_weights=(0.5, 0.34, 0.25), autoupdate=True, dthresh=None, verbose=0, **kwargs):
        """Lambdas that give each target mean weight of the centre pattern, by Nelder-Mead
        on the nearest-neighbour distances (nlpar_cpu.py:94-184); the median becomes
        ``self.lam`` when ``autoupdate``."""
        import scipy.optimize as opt
        target_weights = np.asarray(target_weights)
        if dthresh is not None:
            self.dthresh = dthresh
        dth = np.float64(_scalar(self.dthresh))

        def loptfunc(lam, d2, tw, dthresh):
            temp = np.maximum(d2, dthresh)
            dw = np.exp(-(temp) / lam ** 2)
            w = np.sum(dw, axis=2) + 1e-12
            return np.mean(np.abs(tw - 1.0 / w))

        sigma, d2, n2 = self.calcsigma(nn=1, **kwargs)
        vals = []
        stride = 1 if sigma.size < 1e6 else 2
        for tw in target_weights:
            r = opt.minimize(loptfunc, 1.0, args=(d2[0::stride, 0::stride, :], tw, np.float32(dth)),
                             method='Nelder-Mead', bounds=[[0.001, 10.0]], options={'fatol': 0.0001})
            vals.append(r['x'])
        vals = np.asarray(vals)
        if verbose >= 1:
            print("Range of lambda values: ", vals.flatten())
            print("Optimal Choice: ", np.median(vals))
        if autoupdate:
            self.lam = np.median(vals)
        if self.sigma is None:
            self.sigma = sigma
        self._rows_cache = None
        return vals.flatten()

    # -- NLPAR (nlpar_cpu.py:283-540) ------------------------------------------------------
    def calcnlpar(self, chunksize=0, searchradius=None, lam=None, dthresh=None, saturation_protect=None,
                  automask=None, stem_scale=None, filename=None, fileout=None, reset_sigma=False, backsub=False,
                  rescale=False, verbose=0, diff_offset=None, **kwargs):
        """Average every pattern of the file over its search window and write the result to
        the output file (same type, clipped); returns the output path.  ``chunksize`` > 0
        bounds the device bytes of patterns per band (nlpar_cpu.py:352-358)."""
        if chunksize and chunksize > 0:
            self.band_bytes = int(chunksize)
        if searchradius is not None:
            self.searchradius = searchradius
        sr = int(_scalar(self.searchradius))
        if lam is not None:
            self.lam = lam
        lamv = np.float32(_scalar(self.lam))
        if dthresh is not None:
            self.dthresh = dthresh
        dth = np.float32(_scalar(self.dthresh))
        if diff_offset is not None:
            self.diff_offset = diff_offset
        doff = np.float32(_scalar(self.diff_offset))
        if saturation_protect is not None:
            self.saturation_protect = saturation_protect
        satp = bool(_scalar(self.saturation_protect))
        if automask is not None:
            self.automask = automask
        if stem_scale is not None:
            self.stem_scale = stem_scale
        stem = _scalar(self.stem_scale) is True
        backsub = backsub is True
        if filename is not None:
            self.setfile(filepath=filename)
        f = self.getinfileobj()
        self.setoutfile(f, filepath=fileout)
        fo = self.getoutfileobj()
        dev = self._device()
        torch = _torch()
        nrows, ncols = int(self.nrows), int(self.ncols)
        idx = self._indices(int(f.patternH), int(f.patternW), _scalar(self.automask))
        if reset_sigma:
            self.sigma = None
        if self.sigma is not None and np.asarray(self.sigma).size == 1:
            self.sigma = np.zeros((nrows, ncols), np.float32) + np.asarray(self.sigma).reshape(-1)[0]
        bands = self._bands(f, sr, float_rows=stem or backsub)
        self._options(f, dev, stem, backsub, bands)
        dmin = self._opt[0] if self._opt is not None else None
        if self.sigma is None or np.asarray(self.sigma).shape != (nrows, ncols):
            self.sigma = self.calcsigma(nn=1, saturation_protect=satp, _keep_options=True)[0]
        sigma = np.asarray(self.sigma, np.float32)
        mx = self._scan_max(f, bands, dev)
        odt = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint16): torch.uint16}[np.dtype(fo.filedatatype)]
        # rescale: each averaged pattern stretched to the output type's range (:392-397)
        top = float(np.iinfo(fo.filedatatype).max) if (rescale == True and  # noqa: E712 (as the reference)
                                                       np.issubdtype(fo.filedatatype, np.integer)) else None
        post = stem or top is not None
        newmax, newmin = -np.inf, np.inf
        ws = 2 * sr + 1
        if nrows < ws or ncols < ws:
            raise ValueError(f"the scan ({nrows} x {ncols}) is smaller than the search window ({ws} x {ws})")
        for r0, r1 in bands:
            # the band's rows plus sr on each side, and never fewer than the window: every
            # computed row then finds its whole-scan window inside the band
            a = max(0, min(r0 - sr, nrows - ws))
            b = min(nrows, max(r1 + sr, a + ws))
            x = self._load_rows(f, a, b, dev)
            sg = torch.as_tensor(sigma[a:b], device=dev)
            out = nlpar_device(x, b - a, ncols, idx, sg, lam=lamv, sr=sr, dthresh=dth, saturation_protect=satp,
                               diff_offset=doff, calclim=(0, r0 - a, ncols, r1 - r0),
                               out_dtype=torch.float32 if post else odt, mxval=mx)
            if post:
                out = finish_device(out, stem_min=dmin if stem else None, rescale_top=top, out_dtype=odt)
            host = out.cpu().numpy()
            newmax, newmin = max(newmax, float(host.max())), min(newmin, float(host.min()))
            fo.pat_writer(host, r0 * ncols, (r1 - r0) * ncols, fo.filedatatype)
            if verbose >= 2:
                print("rows complete: ", r1, "/", nrows, sep='', end='\r')
        fo.datamin, fo.datamax = newmin, newmax
        self._rows_cache = None
        self._opt = None
        return str(fo.filepath)

    def auto_nlpar(self, filename=None, fileout=None, lindex=1, **kwargs):
        """opt_lambda, then calcnlpar with the ``lindex``-th lambda (nlpar_cpu.py:81-91)."""
        if filename is not None:
            self.setfile(filename)
        lam = self.opt_lambda(autoupdate=True, **kwargs)
        if 'lam' not in kwargs:
            kwargs['lam'] = lam[int(lindex)]
        return self.calcnlpar(fileout=fileout, **kwargs)

    # reference aliases (nlpar_cpu.py:938-946)
    calcsigma_cpu = calcsigma
    opt_lambda_cpu = opt_lambda
    calcnlpar_cpu = calcnlpar
