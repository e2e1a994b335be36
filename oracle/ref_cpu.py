"""Reference-faithful CPU restatement of PyEBSDIndex's CPU band detection.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  This is the parity target of the
device path and the timed CPU baseline.  It follows, step by step:

* ``pyebsdindex/radon_fast.py:64-193``   plan set-up (same numpy expressions, so the
  same float32/float64 promotions and half-even ``np.round``),
* ``pyebsdindex/radon_fast.py:242-310``  ``radon_faster`` / ``rdn_loops`` (C restatement),
* ``pyebsdindex/band_detect.py:138-294`` ``band_detect_setup`` (rhoMax, rdnmask, kernel, padding),
* ``pyebsdindex/band_detect.py:411-526`` ``find_bands`` chunk driver, normmax, theta/rho,
* ``pyebsdindex/band_detect.py:562-637`` ``rdn_conv`` (in-place padding + ``fftconvolve``),
* ``pyebsdindex/band_detect.py:639-664`` ``rdn_local_max`` (C restatement of ``_grey_dilation``),
* ``pyebsdindex/band_detect.py:667-809`` ``band_label`` (C restatement of ``band_label_numba``).

It never follows ``opencl/band_detect_cl.py`` or ``clkernels.cl``.
"""
import ctypes
import os

import numpy as np
import scipy.fft
import scipy.ndimage as scipyndim
import scipy.signal as scipysignal

from . import build as _build

RADEG = 180.0 / np.pi
DEGRAD = np.pi / 180.0

_LIB = None


def lib():
    """Load (building if needed) the C restatement of the Numba loops."""
    global _LIB
    if _LIB is None:
        path = _build.build(force=False)
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        I = ctypes.c_int64
        L.ora_rdn_loops.argtypes = [P, P, I, I, I, I, I, P, I, I, I, I, ctypes.c_float]
        L.ora_rdn_loops.restype = None
        L.ora_grey_dilation.argtypes = [P, P, I, I, I, I, I]
        L.ora_grey_dilation.restype = None
        L.ora_band_label.argtypes = [I, I, P, P, I, I, P, P, P, P, P, P]
        L.ora_band_label.restype = None
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class RefRadon:
    """Restates ``radon_fast.Radon`` (radon_fast.py:43-312)."""

    def __init__(self, imageDim, nTheta=180, nRho=90, rhoMax=None, mask=None, maskindex=None,
                 missingval=-1.0):
        self.nTheta = nTheta
        self.nRho = nRho
        self.rhoMax = rhoMax
        self.mask = mask
        self.maskindex = maskindex
        self.missingval = missingval
        self.imDim = np.asarray(imageDim, dtype=np.int64)
        self.radon_plan_setup()

    def _masksetup(self):
        # radon_fast.py:168-191, as called from radon_plan_setup (no arguments: a mask
        # given at construction keeps its dtype; the plan truncates it with astype(int64))
        nPx = int(self.imDim[0]) * int(self.imDim[1])
        if self.mask is None:
            self.mask = np.ones(self.imDim)
        if self.maskindex is None:
            self.maskindex = np.arange(nPx, dtype=np.int64).reshape(self.imDim)
        if self.maskindex.max() >= nPx:
            raise Exception("max of index array must be less than the total number of pixel in the array")

    def radon_plan_setup(self):
        # radon_fast.py:79-165, same expressions (and so the same numpy promotions).
        self._masksetup()
        imDim = self.imDim
        if self.rhoMax is None:
            self.rhoMax = np.linalg.norm(imDim) * 0.5
        deltaRho = float(2 * self.rhoMax) / (self.nRho)
        self.theta = np.arange(self.nTheta, dtype=np.float32) * 180.0 / self.nTheta
        self.rho = np.arange(self.nRho, dtype=np.float32) * deltaRho - (self.rhoMax - deltaRho)
        xmin = -1.0 * (self.imDim[1] - 1) * 0.5
        ymin = -1.0 * (self.imDim[0] - 1) * 0.5
        sTheta = np.sin(self.theta * DEGRAD)
        cTheta = np.cos(self.theta * DEGRAD)
        thetatest = np.abs(sTheta) >= (np.sqrt(2.) * 0.5)
        m = np.arange(self.imDim[1], dtype=np.uint32)
        n = np.arange(self.imDim[0], dtype=np.uint32)
        a = -1.0 * np.where(thetatest == 1, cTheta, sTheta)
        a /= np.where(thetatest == 1, sTheta, cTheta)
        b = xmin * cTheta + ymin * sTheta
        outofbounds = self.imDim[0] * self.imDim[1] + 1
        plan = np.zeros([self.nRho, self.nTheta, self.imDim.max()], dtype=np.uint64) + outofbounds
        for i in np.arange(self.nTheta):
            b1 = self.rho - b[i]
            if thetatest[i]:
                b1 /= sTheta[i]
                b1 = b1.reshape(self.nRho, 1)
                indx_y = np.round(a[i] * m + b1).astype(np.int64)
                indx_y = np.where(indx_y < 0, outofbounds, indx_y)
                indx_y = np.where(indx_y >= self.imDim[0], outofbounds, indx_y)
                indx1D = np.clip(m + self.imDim[1] * indx_y, 0, outofbounds)
                plan[:, i, 0:self.imDim[1]] = indx1D
            else:
                b1 /= cTheta[i]
                b1 = b1.reshape(self.nRho, 1)
                indx_x = np.round(a[i] * n + b1).astype(np.int64)
                indx_x = np.where(indx_x < 0, outofbounds, indx_x).astype(np.int64)
                indx_x = np.where(indx_x >= self.imDim[1], outofbounds, indx_x).astype(np.int64)
                indx1D = np.clip(indx_x + self.imDim[1] * n, 0, outofbounds).astype(np.int64)
                plan[:, i, 0:self.imDim[0]] = indx1D
        tempindx = plan.flatten().astype(np.int64)
        mask = np.concatenate((self.mask.flatten().astype(np.int64), np.array([0, 0], dtype=np.int64)))
        tempindx = np.where(mask[tempindx] > 0, tempindx, outofbounds)
        maskindex = np.concatenate((self.maskindex.flatten(), np.array([-1, -1]))).astype(np.int64)
        tempindx = np.where(maskindex[tempindx] >= 0, maskindex[tempindx], outofbounds)
        self.indexPlan = tempindx.reshape([self.nRho, self.nTheta, self.imDim.max()])
        self.indexPlan.sort(axis=-1)

    def radon_faster(self, imageIn, padding=np.array([0, 0]), background=None, normalization=True):
        # radon_fast.py:242-279 (fixArtifacts is never set on the band-detection path)
        shapeIm = np.shape(imageIn)
        nIm = 1 if imageIn.ndim == 2 else shapeIm[0]
        if background is None:
            image = (imageIn.reshape(-1)).astype(np.float32)
        else:
            image = imageIn - background
            image = (image.reshape(-1)).astype(np.float32)
        image = np.ascontiguousarray(image)
        nPx = shapeIm[-1] * shapeIm[-2]
        padding = np.asarray(padding)
        radon = np.full([self.nRho + 2 * padding[0], self.nTheta + 2 * padding[1], nIm],
                        self.missingval, dtype=np.float32)
        plan = np.ascontiguousarray(self.indexPlan, dtype=np.int64)
        lib().ora_rdn_loops(_ptr(image), _ptr(plan), nIm, nPx, self.nRho, self.nTheta,
                            plan.shape[2], _ptr(radon), radon.shape[0], radon.shape[1],
                            int(padding[0]), int(padding[1]), np.float32(normalization > 0))
        return radon


class RefBandDetect:
    """Restates ``band_detect.BandDetect`` (band_detect.py:67-526) for find_bands."""

    def __init__(self, patDim, nTheta=180, nRho=90, tSigma=None, rSigma=None, rhoMaskFrac=0.1,
                 nBands=9, nPhases=10, patternmask=None, patternmaskindex=None):
        self.patDim = np.asarray(patDim, dtype=np.int64)
        self.nTheta = nTheta
        self.nRho = nRho
        self.tSigma = tSigma
        self.rSigma = rSigma
        self.rhoMaskFrac = rhoMaskFrac
        self.nBands = nBands
        self.peakPad = np.array([11, 11])
        self.padding = np.array([11, 11])
        self.backgroundsub = None
        self.dataType = np.dtype([('id', np.int32), ('max', np.float32), ('normmax', np.float32),
                                  ('maxloc', np.float32, (2)), ('avemax', np.float32),
                                  ('aveloc', np.float32, (2)), ('pqmax', np.float32),
                                  ('width', np.float32), ('theta', np.float32), ('rho', np.float32),
                                  ('valid', np.int8), ('band_match_index', np.int64, (nPhases,)),
                                  ('bandfit', np.float64, (nPhases,))])
        self._setup(patternmask, patternmaskindex)

    def _setup(self, patternmask, patternmaskindex):
        # band_detect.py:185-290
        if (self.rhoMaskFrac < 1) and (self.rhoMaskFrac > 0):
            self.rhoMax = 0.5 * np.float32(self.patDim.min())
        else:
            self.rhoMax = 0.5 * np.float32(np.sqrt(np.float32(self.patDim[-2]) ** 2 + np.float32(self.patDim[-1]) ** 2))
        self.dTheta = 180.0 / self.nTheta
        self.dRho = self.rhoMax / np.float32(self.nRho)
        self.radonPlan = RefRadon(self.patDim, nTheta=self.nTheta, nRho=self.nRho, rhoMax=self.rhoMax,
                                  mask=patternmask, maskindex=patternmaskindex)
        back = np.array(self.radonPlan.mask > 0).astype(np.float32)
        rdnmask = self.radonPlan.radon_faster(back, normalization=False)
        rdnmask = rdnmask > 0
        rdnmask = rdnmask.squeeze()
        if self.rhoMaskFrac >= (1 - 1.0e-12):
            s = np.ones((3, 1))
            rdnmask = scipyndim.binary_erosion(rdnmask, structure=s, iterations=int(self.rhoMaskFrac))
        else:
            cmask = self._circmask(back)
            rdncmask = rdnmask.copy()
            if self.rhoMaskFrac > 0:
                mskinx = int(self.nRho * self.rhoMaskFrac)
                if mskinx == 0:
                    mskinx = 1
                rdncmask[0:mskinx, :] = 0
                rdncmask[-mskinx:, :] = 0
            elif self.rhoMaskFrac < 0:
                cmask = (cmask < self.rhoMax * (1.0 + self.rhoMaskFrac)).astype(np.float32)
                rdncmask = (self.radonPlan.radon_faster(cmask) > 0).squeeze()
            rdnmask *= (rdncmask > 0).astype(bool)
        self.rdnmask = rdnmask > 0
        if self.tSigma is None:
            self.tSigma = 1.0 / self.dTheta
        if self.rSigma is None:
            self.rSigma = 0.5 / np.float32(self.dRho)
        ksz = np.array([np.max([np.int64(4 * self.rSigma), 5]), np.max([np.int64(4 * self.tSigma), 5])])
        ksz = ksz + ((ksz % 2) == 0)
        kernel = np.zeros(ksz, dtype=np.float32)
        kernel[(ksz[0] / 2).astype(int), (ksz[1] / 2).astype(int)] = 1
        kernel = -1.0 * scipyndim.gaussian_filter(kernel, [self.rSigma, self.tSigma], order=[2, 0])
        kernel *= 1.0 / np.max(kernel)
        self.kernel = kernel.reshape((1, ksz[0], ksz[1]))
        self.peakPad = np.array(np.around([2 * ksz[0], 2 * ksz[1]]), dtype=np.int64)
        self.peakPad += 1 - np.mod(self.peakPad, 2)
        self.padding = np.array([np.max([self.peakPad[0], self.padding[0]]),
                                 np.max([self.peakPad[1], self.padding[1]])])

    @staticmethod
    def _circmask(im):
        # band_detect.py:918-931
        nx = im.shape[-1]
        ny = im.shape[-2]
        cx = (np.arange(nx) - nx * 0.5) ** 2
        cy = (np.arange(ny) - ny * 0.5) ** 2
        return np.sqrt(np.broadcast_to(cx.reshape(1, nx), (ny, nx)) + np.broadcast_to(cy, (nx, ny)).T)

    # ---- per-chunk stages -------------------------------------------------------------
    def rdn_conv(self, radon):
        # band_detect.py:562-637 (padding in place, then fftconvolve with all host cores)
        p0, p1 = int(self.padding[0]), int(self.padding[1])
        shp = radon.shape
        if p1 > 0:
            radon[:, 0:p1, :] = np.flip(radon[:, -2 * p1:-p1, :], axis=0)
            radon[:, -p1:, :] = np.flip(radon[:, p1:2 * p1, :], axis=0)
        if p0 > 0:
            radon[0:p0, :, :] = radon[p0, :, :].reshape(1, shp[1], shp[2])
            radon[-p0:, :, :] = radon[-p0 - 1, :, :].reshape(1, shp[1], shp[2])
        k = np.copy(self.kernel[0, :, :])[:, :, np.newaxis]
        with scipy.fft.set_workers(os.cpu_count()):
            rdnConv = scipysignal.fftconvolve(radon, k, mode='same')
        sl = rdnConv[p0:shp[1] - p0, p1:shp[1] - p1, :]
        mns = sl.min(axis=0).min(axis=0)
        mx = sl.max(axis=0).max(axis=0)
        ave = np.mean(sl, axis=(0, 1))
        return rdnConv, [mns, ave, mx]

    def rdn_local_max(self, rdn):
        # band_detect.py:639-664
        rdn = np.ascontiguousarray(rdn, dtype=np.float32)
        n0, n1, nq = rdn.shape
        tmp = np.empty_like(rdn)
        out = np.empty_like(rdn)
        L = lib()
        L.ora_grey_dilation(_ptr(rdn), _ptr(tmp), n0, n1, nq, 1, int(self.peakPad[1]))
        L.ora_grey_dilation(_ptr(tmp), _ptr(out), n0, n1, nq, int(self.peakPad[0]), 1)
        lmax = out == rdn
        p0, p1 = int(self.padding[0]), int(self.padding[1])
        maskrnd = np.zeros((self.nRho + 2 * p0, self.nTheta + 2 * p1, 1), dtype=np.ubyte)
        maskrnd[p0:-p0, p1:-p1, 0] = self.rdnmask.astype(np.ubyte)
        lmax *= maskrnd.astype(bool)
        return lmax

    def band_label(self, nPats, rdnConv, lmax):
        # band_detect.py:667-689
        nB = self.nBands
        bmax = np.zeros((nPats, nB), dtype=np.float32) - 2.0e6
        bavemax = np.zeros((nPats, nB), dtype=np.float32) - 2.0e6
        bvalid = np.zeros((nPats, nB), dtype=np.int8)
        bmaxloc = np.zeros((nPats, nB, 2), dtype=np.float32)
        baveloc = np.zeros((nPats, nB, 2), dtype=np.float32)
        bwidth = np.zeros((nPats, nB), dtype=np.float32)
        conv = np.ascontiguousarray(rdnConv, dtype=np.float32)
        lm = np.ascontiguousarray(lmax, dtype=np.uint8)
        lib().ora_band_label(nB, nPats, _ptr(conv), _ptr(lm), conv.shape[0], conv.shape[1],
                             _ptr(bmax), _ptr(bavemax), _ptr(bmaxloc), _ptr(baveloc), _ptr(bvalid),
                             _ptr(bwidth))
        bandData = np.zeros((nPats, nB), dtype=self.dataType)
        bandData['max'] = bmax
        bandData['normmax'] = bmax
        bandData['avemax'] = bavemax
        bandData['maxloc'] = bmaxloc
        bandData['aveloc'] = baveloc
        bandData['valid'] = bvalid
        bandData['width'] = bwidth
        bandData['maxloc'] -= self.padding.reshape(1, 1, 2)
        bandData['aveloc'] -= self.padding.reshape(1, 1, 2)
        return bandData

    def find_bands(self, patternsIn, chunksize=512, keep=None):
        """band_detect.py:411-526.  ``keep`` (a dict) receives the last chunk's Radon,
        convolved Radon, stats and local-max mask for stage-by-stage comparisons."""
        patterns = np.expand_dims(patternsIn, axis=0) if patternsIn.ndim == 2 else patternsIn
        nPats = patterns.shape[0]
        bandData = np.zeros((nPats, self.nBands), dtype=self.dataType)
        bandData['band_match_index'] = -100
        if chunksize <= 0:
            chunk_start_end = [[0, nPats]]
        else:
            nchunks = (np.ceil(nPats / chunksize)).astype(np.int64)
            chunk_start_end = [[i * chunksize, (i + 1) * chunksize] for i in range(nchunks)]
            chunk_start_end[-1][1] = nPats
        for chnk in chunk_start_end:
            rdnNorm = self.radonPlan.radon_faster(patterns[chnk[0]:chnk[1], :, :], self.padding,
                                                  background=self.backgroundsub)
            if keep is not None:
                keep['radon'] = rdnNorm.copy()
            rdnConv, imageminavemax = self.rdn_conv(rdnNorm)
            lMaxRdn = self.rdn_local_max(rdnConv)
            bandDataChunk = self.band_label(chnk[1] - chnk[0], rdnConv, lMaxRdn)
            bndnorm = bandDataChunk['normmax']
            bndnorm -= imageminavemax[0].reshape(chnk[1] - chnk[0], 1)
            bndnorm /= (imageminavemax[1] - imageminavemax[0]).reshape(chnk[1] - chnk[0], 1).clip(1e-7)
            bandDataChunk['normmax'] = bndnorm
            bandData[chnk[0]:chnk[1]] = bandDataChunk
            if keep is not None:
                keep['radon_padded'] = rdnNorm
                keep['conv'] = rdnConv
                keep['stats'] = imageminavemax
                keep['lmax'] = lMaxRdn
        theta = np.pi - np.interp(bandData['aveloc'][:, :, 1], np.arange(self.radonPlan.nTheta),
                                  self.radonPlan.theta) / RADEG
        rho = -1.0 * np.interp(bandData['aveloc'][:, :, 0], np.arange(self.radonPlan.nRho),
                               self.radonPlan.rho)
        bandData['theta'][:] = theta
        bandData['rho'][:] = rho
        return bandData
