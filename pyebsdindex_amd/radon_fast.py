"""Radon plan (host set-up) and the gnomonic band -> plane-normal map.

Mirrors ``pyebsdindex.radon_fast.Radon`` (pyebsdindex/radon_fast.py:43-397):

* ``radon_plan_setup`` builds the same ``indexPlan`` as radon_fast.py:64-165 -- the
  float32 theta/sin/cos and rho grid, float64 line offsets, half-even ``np.round`` --
  but vectorised over all theta at once instead of the reference's per-theta loop.
  It runs once per pattern size on the host; the device never recomputes it (the
  float32 ``np.sin``/``np.cos`` values are numpy's, so the plan is built by numpy).
* ``radon_faster`` runs the Radon gather on the GPU (``ebsd_radon``).
* ``radon2pole`` is the reference's host-side numpy map (radon_fast.py:314-397).
"""
import numpy as np

RADEG = 180.0 / np.pi
DEGRAD = np.pi / 180.0


class Radon:
    def __init__(self, image=None, imageDim=None, nTheta=180, nRho=90, rhoMax=None,
                 mask=None, maskindex=None, missingval=-1.0):
        self.nTheta = nTheta
        self.nRho = nRho
        self.rhoMax = rhoMax
        self.mask = mask
        self.maskindex = maskindex
        self.missingval = missingval
        self.theta = None
        self.rho = None
        self.imDim = None
        self.indexPlan = None
        if image is not None:
            self.imDim = np.asarray(image.shape[-2:])
            self.radon_plan_setup()
        elif imageDim is not None:
            self.imDim = np.asarray(imageDim[-2:])
            self.radon_plan_setup()

    # -- set-up (radon_fast.py:64-193) --------------------------------------------------
    def _masksetup(self, mask=None, maskindex=None):
        if mask is not None:
            self.mask = np.array(mask).astype(int)
        if maskindex is not None:
            self.maskindex = np.array(maskindex).astype(np.int64)
        if self.imDim is None:
            return
        nPx = int(self.imDim[0]) * int(self.imDim[1])
        if self.mask is None:
            self.mask = np.ones(self.imDim)
        if self.maskindex is None:
            self.maskindex = np.arange(nPx, dtype=np.int64).reshape(self.imDim)
        if np.max(self.maskindex) >= nPx:
            raise Exception("max of index array must be less than the total number of pixel in the array")

    def radon_plan_setup(self, image=None, imageDim=None, nTheta=None, nRho=None, rhoMax=None,
                         mask=None, maskindex=None):
        if image is None and imageDim is not None:
            self.imDim = np.asarray(imageDim, dtype=np.int64)
        elif image is not None:
            self.imDim = np.asarray(np.shape(image)[-2:])
        if self.imDim is None:
            return -1
        if mask is not None:
            self.mask = mask
        if maskindex is not None:
            self.maskindex = maskindex
        self._masksetup()
        if nTheta is not None:
            self.nTheta = nTheta
        if nRho is not None:
            self.nRho = nRho
        if rhoMax is not None:
            self.rhoMax = rhoMax
        if self.rhoMax is None:
            self.rhoMax = np.linalg.norm(self.imDim) * 0.5

        H, W = int(self.imDim[0]), int(self.imDim[1])
        # grids: the same expressions as the reference, so the same float32 values
        dRho = float(2 * self.rhoMax) / self.nRho
        self.theta = np.arange(self.nTheta, dtype=np.float32) * 180.0 / self.nTheta
        self.rho = np.arange(self.nRho, dtype=np.float32) * dRho - (self.rhoMax - dRho)
        s = np.sin(self.theta * DEGRAD)
        c = np.cos(self.theta * DEGRAD)
        steep = np.abs(s) >= (np.sqrt(2.) * 0.5)
        num = np.where(steep, c, s)
        den = np.where(steep, s, c)
        slope = ((-1.0 * num) / den).astype(np.float64)             # float32 ratio, exact in f64
        off0 = (-(W - 1) * 0.5) * c.astype(np.float64) + (-(H - 1) * 0.5) * s.astype(np.float64)
        off = (self.rho.astype(np.float64)[:, None] - off0[None, :]) / den.astype(np.float64)[None, :]

        L = max(H, W)
        nPx = H * W
        sentinel = nPx + 1
        step = np.arange(L, dtype=np.float64)
        coord = np.round(slope[None, :, None] * step[None, None, :] + off[:, :, None]).astype(np.int64)
        k = np.arange(L, dtype=np.int64)[None, None, :]
        st = steep[None, :, None]
        # steep lines step over columns (x = k, y = coord), the others over rows
        in_range = np.where(st, (k < W) & (coord >= 0) & (coord < H),
                            (k < H) & (coord >= 0) & (coord < W))
        pix = np.where(st, k + W * coord, coord + W * k)
        pix = np.where(in_range, pix, sentinel)
        # pattern mask / mask index remap (radon_fast.py:159-163)
        mflat = np.concatenate((np.asarray(self.mask).reshape(-1).astype(np.int64) > 0, [False, False]))
        mi = np.concatenate((np.asarray(self.maskindex).reshape(-1).astype(np.int64), [-1, -1]))
        pix = np.where(mflat[pix], pix, sentinel)
        pix = np.where(mi[pix] >= 0, mi[pix], sentinel)
        pix.sort(axis=-1)
        self.indexPlan = pix
        self.__dict__.pop("_device_plans", None)   # device plans of the previous geometry

    def __getstate__(self):
        """Pickled (as Ray's object store does with the indexer) without device plans: they
        hold library handles and device memory of this process."""
        state = self.__dict__.copy()
        state.pop("_device_plans", None)
        return state

    # -- Radon on the GPU --------------------------------------------------------------
    def radon_faster(self, imageIn, padding=np.array([0, 0]), fixArtifacts=False, background=None,
                     normalization=True):
        """Reference signature (radon_fast.py:242-279); result (nRp, nTp, nIm) float32.

        Runs ``ebsd_radon`` on the current torch device.  Only the normalised Radon of the
        band-detection path is provided on the device."""
        if not normalization:
            raise ValueError("radon_faster: only normalization=True runs on the device")
        from .band_detect import _radon_only_plan
        import torch
        img = np.asarray(imageIn)
        if img.ndim == 2:
            img = img[None]
        if background is not None:
            img = (img - background).astype(np.float32)
        plan = _radon_only_plan(self)
        dev = torch.device("cuda", torch.cuda.current_device())
        t = torch.from_numpy(np.ascontiguousarray(img)).to(dev)
        rad = plan.radon(t)
        rad = rad.permute(1, 2, 0).cpu().numpy()
        pad = np.asarray(padding)
        out = np.full((self.nRho + 2 * pad[0], self.nTheta + 2 * pad[1], img.shape[0]),
                      self.missingval, dtype=np.float32)
        out[pad[0]:pad[0] + self.nRho, pad[1]:pad[1] + self.nTheta, :] = rad
        if fixArtifacts:
            shp = out.shape
            out[:, pad[1], :] = out[:, pad[1] + 1, :]
            out[:, shp[1] - 1 - pad[1], :] = out[:, shp[1] - pad[1] - 2, :]
        return out

    # -- band (theta, rho) -> plane normal (radon_fast.py:314-397) -----------------------
    def radon2pole(self, bandData, PC=None, vendor='EDAX'):
        if PC is None:
            PC = np.array([0.471659, 0.675044, 0.630139])
        ven = str.upper(vendor)
        nPats, nBands = bandData.shape[0], bandData.shape[1]
        theta = bandData['theta'][:]
        rho = bandData['rho'][:]
        st = np.sin(theta)
        ct = np.cos(theta)
        pc = np.asarray(PC, dtype=np.float32).copy()
        if ven != 'EMSOFT':
            if pc.ndim < 2:
                pc = np.tile(pc, nPats).reshape(nPats, 3)
            elif pc.shape[0] != nPats:
                pc = np.tile(pc[0, :], nPats).reshape(nPats, 3)
            t = pc
        else:
            if pc.ndim < 2:
                pc = np.tile(pc, nPats).reshape(nPats, 4)
            elif pc.shape[0] != nPats:
                pc = np.tile(pc[0, :], nPats).reshape(nPats, 4)
            t = pc[:, 0:3]
            t[:, 2] /= pc[:, 3]
        dimf = np.array(self.imDim, dtype=np.float32)
        if ven in ['EDAX']:
            t *= np.array([dimf[1], dimf[0], -np.min(dimf[0:2])])
        if ven in ['OXFORD']:
            t *= np.array([dimf[1], dimf[1], -dimf[1]])
        if ven == 'EMSOFT':
            t[:, 0] *= -1.0
            t += np.array([dimf[1] / 2.0, dimf[0] / 2.0, 0.0])
            t[:, 2] *= -1.0
        if ven in ['KIKUCHIPY', 'BRUKER']:
            t *= np.array([dimf[1], dimf[0], -dimf[0]])
            t[:, 1] = dimf[0] - t[:, 1]
        t = np.broadcast_to(t.reshape(nPats, 1, 3), (nPats, nBands, 3))
        # r = in-plane band direction, p = foot of the perpendicular from the pattern centre
        r = np.zeros((nPats, nBands, 3), dtype=np.float32)
        r[:, :, 0] = -1 * st
        r[:, :, 1] = ct
        p = np.zeros((nPats, nBands, 3), dtype=np.float32)
        p[:, :, 0] = rho * ct + dimf[1] * 0.5
        p[:, :, 1] = rho * st + dimf[0] * 0.5
        n = np.cross(r.reshape(-1, 3), (p - t).reshape(-1, 3))
        nrm = np.linalg.norm(n, axis=1).clip(1e-12)
        n /= nrm.reshape(-1, 1)
        return n.reshape(nPats, nBands, 3)
