"""Drop-in ``BandDetect`` whose ``find_bands`` runs on one MI355X through the C-ABI library.

Mirrors the reference API (pyebsdindex/band_detect.py:67-526 and the GPU subclass
pyebsdindex/opencl/band_detect_cl.py:45-199): the same constructor keywords and
attributes (``nTheta``, ``nRho``, ``tSigma``, ``rSigma``, ``rhoMaskFrac``, ``nBands``,
``padding``, ``peakPad``, ``kernel``, ``rdnmask``, ``radonPlan``, ``dataType``,
``backgroundsub``, ``useCPU``), ``band_detect_setup``, ``collect_background`` and
``find_bands(patterns, verbose, chunksize, clparams, gpu_id)`` returning the same
structured array.  Set-up (plan, masks, kernel) stays on the host as in the reference;
everything inside ``find_bands`` runs on the device.

Differences from the reference, by design:
* no CPU fallback: a device or library error raises (the OpenCL subclass prints and
  re-runs on the CPU, band_detect_cl.py:195-197);
* ties between exactly equal peak values are ordered as a stable argsort would order
  them (Numba's unstable quicksort order, band_detect.py:734, is not reproducible);
* ``background_mode='fast'`` (default) subtracts the background as the float64 line sum
  of the background from the exact integer line sum; ``'exact'`` stages float32(x - b)
  per pixel and sums in float64 exactly as radon_fast.py:255-259,288-310.
"""
from timeit import default_timer as timer

import numpy as np
import scipy.ndimage as scipyndim
import scipy.optimize as scipyopt

from . import _hip
from .radon_fast import Radon, RADEG


def _dtype_code(dt):
    if dt == np.uint8:
        return _hip.EBSD_U8
    if dt == np.uint16:
        return _hip.EBSD_U16
    if dt == np.float32:
        return _hip.EBSD_F32
    return None


def record_phases(dtype):
    """The phase count if ``dtype`` (a ``BandDetect.dataType``) has the packed layout that
    ``ebsd_pack_band_records`` writes (band_detect.py:110-115), else None."""
    dtype = np.dtype(dtype)
    if dtype.names is None or "band_match_index" not in dtype.names:
        return None
    nph = int(np.prod(dtype["band_match_index"].shape))
    f4, i4, i1, i8, f8 = (np.dtype(t) for t in ("<f4", "<i4", "i1", "<i8", "<f8"))
    want = [("id", i4, (), 0), ("max", f4, (), 4), ("normmax", f4, (), 8), ("maxloc", f4, (2,), 12),
            ("avemax", f4, (), 20), ("aveloc", f4, (2,), 24), ("pqmax", f4, (), 32), ("width", f4, (), 36),
            ("theta", f4, (), 40), ("rho", f4, (), 44), ("valid", i1, (), 48),
            ("band_match_index", i8, (nph,), 49), ("bandfit", f8, (nph,), 49 + 8 * nph)]
    if list(dtype.names) != [w[0] for w in want] or dtype.itemsize != 49 + 16 * nph:
        return None
    for name, base, shape, off in want:
        ft, fo = dtype.fields[name][:2]
        if fo != off or ft.base != base or tuple(ft.shape) != tuple(shape):
            return None
    return nph


def auto_pass_size(N, cells, cus, budget_bytes):
    """Patterns per Radon+peak pass when the caller names none: all N while their float32
    Radons (``cells`` each) fit ``budget_bytes``; else the largest whole number of gather
    rounds (4 patterns per CU, ``cus`` CUs) that fits, or what fits where not even one
    round does."""
    N = max(1, int(N))
    cap = max(1, int(budget_bytes) // (int(cells) * 4))
    if N <= cap:
        return N
    unit = 4 * max(1, int(cus))
    return cap // unit * unit if cap >= unit else cap


class DevicePlan:
    """The compiled device plan of one BandDetect on one GPU (``ebsd_plan*``)."""

    def __init__(self, bd, device_index):
        import torch
        self.lib = _hip.load()
        self.device = torch.device("cuda", device_index)
        self.nB = int(bd.nBands)
        self.nR, self.nT = int(bd.nRho), int(bd.nTheta)
        self.H, self.W = int(bd.patDim[-2]), int(bd.patDim[-1])
        self.padding = np.asarray(bd.padding, dtype=np.int64)
        plan = np.ascontiguousarray(bd.radonPlan.indexPlan, dtype=np.int64)
        mask = np.ascontiguousarray(bd.rdnmask, dtype=np.uint8)
        kern = np.ascontiguousarray(bd.kernel[0], dtype=np.float32)
        theta = np.ascontiguousarray(bd.radonPlan.theta, dtype=np.float32)
        rho = np.ascontiguousarray(bd.radonPlan.rho, dtype=np.float32)
        handle = _hip.ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _hip.check(self.lib.ebsd_plan_create(
                self.H, self.W, self.nT, self.nR, plan.ctypes.data, plan.shape[2],
                mask.ctypes.data, kern.ctypes.data, kern.shape[0], kern.shape[1],
                int(bd.padding[0]), int(bd.padding[1]), int(bd.peakPad[0]), int(bd.peakPad[1]),
                self.nB, theta.ctypes.data, rho.ctypes.data, _hip.ctypes.byref(handle)))
        self.handle = handle
        self.kernel_shape = (int(kern.shape[0]), int(kern.shape[1]))
        self.peak_half = (int(bd.peakPad[0]) // 2, int(bd.peakPad[1]) // 2)
        self.peak_kind = 0
        self.bg_key = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.ebsd_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    def info(self):
        out = (_hip.ctypes.c_int64 * 10)()
        _hip.check(self.lib.ebsd_plan_info(self.handle, out))
        keys = ("cells", "pairs", "batches", "peak_tile", "peak_tiles", "peak_lds", "bg_mode", "cus", "slots",
                "line_unit")
        return dict(zip(keys, list(out)))

    def set_background(self, bg, exact):
        key = None if bg is None else (bool(exact), np.asarray(bg, dtype=np.float64).tobytes())
        if key == self.bg_key:
            return
        if bg is None:
            _hip.check(self.lib.ebsd_plan_set_background(self.handle, None, 0))
        else:
            b = np.ascontiguousarray(bg, dtype=np.float64).reshape(-1)
            if b.size != self.H * self.W:
                raise ValueError("background does not match the pattern size")
            _hip.check(self.lib.ebsd_plan_set_background(self.handle, b.ctypes.data, int(bool(exact))))
        self.bg_key = key

    def set_tuning(self, gather_pack=0, gather_threads=0, peak_tile=0, peak_threads=0):
        _hip.check(self.lib.ebsd_plan_set_tuning(self.handle, int(gather_pack), int(gather_threads),
                                                 int(peak_tile), int(peak_threads)))

    def set_stage_pitch(self, pitch):
        """LDS words per staged pattern row in the Radon gather (0 = the pattern width)."""
        _hip.check(self.lib.ebsd_plan_set_stage_pitch(self.handle, int(pitch)))

    def set_row_segments(self, segments):
        """Row segments of the integer Radon gather (0 = automatic, 1..8); identical results."""
        _hip.check(self.lib.ebsd_plan_set_row_segments(self.handle, int(segments)))

    def set_line_unit(self, unit):
        """Gather line order in units of ``unit`` neighbouring cells (0 = default 8);
        identical results, fewer partial output sectors for larger units."""
        _hip.check(self.lib.ebsd_plan_set_line_unit(self.handle, int(unit)))

    def set_peak_kernel(self, kind):
        """0 automatic, 1 theta-tiled work-group kernel, 2 streamed wavefront kernel."""
        _hip.check(self.lib.ebsd_plan_set_peak_kernel(self.handle, int(kind)))
        self.peak_kind = int(kind)

    def set_peak_block(self, theta_cols):
        """Theta columns per wavefront of the streamed peak kernel (0 = default); the
        same results for every value."""
        _hip.check(self.lib.ebsd_plan_set_peak_block(self.handle, int(theta_cols)))

    def stream_block(self):
        """(default theta block, largest theta block) of the streamed peak kernel, or None
        where it does not apply (host restatement of the plan's rule)."""
        kr, kt = int(self.kernel_shape[0]), int(self.kernel_shape[1])
        hr, ht = int(self.peak_half[0]), int(self.peak_half[1])
        if not (hr == kr and ht == kt and (kr, kt) in ((9, 5), (5, 5))):
            return None
        usable = 64 - 2 * (kt // 2 + ht)
        nblk = -(-self.nT // usable)
        if nblk > 16:
            return None
        return -(-self.nT // nblk), usable

    # -- device entry points --------------------------------------------------------------
    def alloc_out(self, N):
        import torch
        f = dict(dtype=torch.float32, device=self.device)
        return {
            "max": torch.empty((N, self.nB), **f), "normmax": torch.empty((N, self.nB), **f),
            "avemax": torch.empty((N, self.nB), **f), "maxloc": torch.empty((N, self.nB, 2), **f),
            "aveloc": torch.empty((N, self.nB, 2), **f), "width": torch.empty((N, self.nB), **f),
            "theta": torch.empty((N, self.nB), **f), "rho": torch.empty((N, self.nB), **f),
            "valid": torch.empty((N, self.nB), dtype=torch.int8, device=self.device),
            "stats": torch.empty((N, 3), **f),
        }

    @staticmethod
    def bands_struct(out):
        b = _hip.EbsdBands()
        for k in ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid", "stats"):
            setattr(b, "d_" + k, out[k].data_ptr())
        return b

    def pack_records(self, out, n_phases, dst=None, stream=None):
        """``ebsd_pack_band_records``: the bands of ``out`` (the dict ``find_bands`` returns,
        N patterns) as N*nB records in the byte layout of ``BandDetect.dataType`` with
        ``n_phases`` phases, into ``dst`` (a device uint8 tensor of N*nB*(49+16*n_phases)
        bytes; allocated here unless given).  Returns ``dst``."""
        import torch
        N = int(out["max"].shape[0])
        nbytes = N * self.nB * (49 + 16 * int(n_phases))
        if dst is None:
            dst = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        if dst.dtype != torch.uint8 or dst.device != self.device or dst.numel() < nbytes or not dst.is_contiguous():
            raise ValueError(f"dst must be a contiguous uint8 tensor of at least {nbytes} bytes on {self.device}")
        s = _hip.launch_stream(stream, self.device)
        b = self.bands_struct(out)
        _hip.check(self.lib.ebsd_pack_band_records(_hip.ctypes.byref(b), N, self.nB, int(n_phases),
                                                   dst.data_ptr(), s.cuda_stream))
        _hip.keep_for_stream(s, dst, *out.values())
        return dst

    def _check_pats(self, pats):
        import torch
        if not isinstance(pats, torch.Tensor) or pats.device != self.device:
            raise ValueError(f"patterns must be a torch tensor on {self.device}")
        if pats.dim() == 2:
            pats = pats.unsqueeze(0)
        if pats.dim() != 3 or pats.shape[1] != self.H or pats.shape[2] != self.W:
            raise ValueError(f"patterns must be (N, {self.H}, {self.W})")
        pats = pats.contiguous()
        code = {torch.uint8: _hip.EBSD_U8, torch.uint16: _hip.EBSD_U16,
                torch.float32: _hip.EBSD_F32}.get(pats.dtype)
        return pats, code

    # Radon workspace of one pass when the caller names no chunk
    AUTO_WORKSPACE_BYTES = 1 << 30

    def auto_chunk(self, N):
        """Patterns per Radon+peak pass when the caller names none: all N while their
        Radons fit a 1 GiB workspace, else whole gather rounds (4 patterns per CU) up to
        that size.  A large pass fills the GPU better at its ends: the peak kernel's
        4-wavefront work-groups take 1.6 rounds of the 1,280 resident slots at 2,048
        patterns, and one 16,384-pattern pass at 120x120 runs 5% faster per pattern than
        eight passes of 2,048."""
        return auto_pass_size(N, self.nR * self.nT, int(self.info()["cus"]), self.AUTO_WORKSPACE_BYTES)

    def workspace(self, N, chunk=None):
        import torch
        if chunk is None:
            chunk = self.auto_chunk(N)
        n = max(1, min(int(N), int(chunk)))
        nbytes = int(self.lib.ebsd_workspace_bytes(self.handle, n))
        return torch.empty(nbytes // 4, dtype=torch.float32, device=self.device)

    def find_bands(self, pats, out=None, work=None, chunk=None, stream=None, flags=0):
        """``ebsd_find_bands`` on device patterns; returns the dict of device outputs.
        ``chunk``: patterns per Radon+peak pass (None: ``auto_chunk``); a given ``work``
        sets it instead (its size in Radons)."""
        import torch
        pats, code = self._check_pats(pats)
        if code is None:
            raise TypeError(f"unsupported device pattern dtype {pats.dtype}")
        N = pats.shape[0]
        if out is None:
            out = self.alloc_out(N)
        if work is None:
            work = self.workspace(N, chunk)
        s = _hip.launch_stream(stream, self.device)
        b = self.bands_struct(out)
        _hip.check(self.lib.ebsd_find_bands(self.handle, pats.data_ptr(), code | flags, N,
                                            _hip.ctypes.byref(b), work.data_ptr(),
                                            work.numel() * 4, s.cuda_stream))
        _hip.keep_for_stream(s, pats, work, *out.values())
        return out

    def find_bands_noalloc(self, pats, out=None, stream=None, flags=0):
        """``ebsd_find_bands_noalloc`` on device patterns: the library allocates no device
        memory (the chunks' Radons live in its per-GPU scratch array); ``out`` is allocated
        here unless given.  Calls that overlap on one GPU race on that array."""
        pats, code = self._check_pats(pats)
        if code is None:
            raise TypeError(f"unsupported device pattern dtype {pats.dtype}")
        N = pats.shape[0]
        if out is None:
            out = self.alloc_out(N)
        s = _hip.launch_stream(stream, self.device)
        b = self.bands_struct(out)
        _hip.check(self.lib.ebsd_find_bands_noalloc(self.handle, pats.data_ptr(), code | flags, N,
                                                    _hip.ctypes.byref(b), s.cuda_stream))
        _hip.keep_for_stream(s, pats, *out.values())
        return out

    def radon(self, pats, stream=None, flags=0, out=None):
        import torch
        pats, code = self._check_pats(pats)
        if code is None:
            raise TypeError(f"unsupported device pattern dtype {pats.dtype}")
        N = pats.shape[0]
        rad = out if out is not None else torch.empty((N, self.nR, self.nT), dtype=torch.float32,
                                                      device=self.device)
        s = _hip.launch_stream(stream, self.device)
        _hip.check(self.lib.ebsd_radon(self.handle, pats.data_ptr(), code | flags, N, rad.data_ptr(),
                                       s.cuda_stream))
        _hip.keep_for_stream(s, pats, rad)
        return rad

    def peaks(self, radon, out=None, stream=None):
        import torch
        radon = radon.contiguous()
        N = radon.shape[0]
        if out is None:
            out = self.alloc_out(N)
        s = _hip.launch_stream(stream, self.device)
        b = self.bands_struct(out)
        _hip.check(self.lib.ebsd_peaks(self.handle, radon.data_ptr(), N, _hip.ctypes.byref(b),
                                       s.cuda_stream))
        _hip.keep_for_stream(s, radon, *out.values())
        return out

    def conv(self, radon, stream=None):
        import torch
        radon = radon.contiguous()
        N = radon.shape[0]
        p0, p1 = int(self.padding[0]), int(self.padding[1])
        c = torch.empty((N, self.nR + 2 * p0, self.nT + 2 * p1), dtype=torch.float32, device=self.device)
        s = _hip.launch_stream(stream, self.device)
        _hip.check(self.lib.ebsd_conv(self.handle, radon.data_ptr(), N, c.data_ptr(), s.cuda_stream))
        _hip.keep_for_stream(s, radon, c)
        return c


def _radon_only_plan(radon):
    """A device plan for ``Radon.radon_faster`` (no peak stage configured)."""
    import torch

    class _Shim:
        pass
    key = ("_dev_plan", torch.cuda.current_device())
    cache = radon.__dict__.setdefault("_device_plans", {})
    if key in cache:
        return cache[key]
    s = _Shim()
    s.nBands, s.nRho, s.nTheta = 1, radon.nRho, radon.nTheta
    s.patDim = np.asarray(radon.imDim)
    s.padding = np.array([1, 1])
    s.peakPad = np.array([3, 3])
    s.radonPlan = radon
    s.rdnmask = np.zeros((radon.nRho, radon.nTheta), dtype=np.uint8)
    s.kernel = np.ones((1, 1, 1), dtype=np.float32)
    plan = DevicePlan(s, torch.cuda.current_device())
    cache[key] = plan
    return plan


class HipDevice:
    """Device descriptor returned by ``getopenclparam`` (role of opencl/openclparam.py:62-129)."""

    def __init__(self, index):
        import torch
        p = torch.cuda.get_device_properties(index)
        self.index = index
        self.name = p.name
        self.global_mem_size = int(p.total_memory)
        self.max_compute_units = int(p.multi_processor_count)


class HipParam:
    def __init__(self, gpu_id=None):
        import torch
        n = torch.cuda.device_count() if torch.cuda.is_available() else 0
        self.gpu = [HipDevice(i) for i in range(n)]
        self.ngpu = len(self.gpu)
        self.gpu_id = 0 if gpu_id is None else int(gpu_id)
        self.gpusharedmem = False

    def get_queue(self, gpu_id=None, **kwargs):
        if gpu_id is not None:
            self.gpu_id = int(gpu_id)
        return self


def getopenclparam(**kwargs):
    """Device factory with the reference's name (band_detect_cl.py:719-722); None without a GPU."""
    try:
        import torch
        if not torch.cuda.is_available():
            return None
    except Exception:
        return None
    return HipParam(**kwargs)


class BandDetect:
    def __init__(self, patterns=None, patDim=None, nTheta=180, nRho=90, tSigma=None, rSigma=None,
                 rhoMaskFrac=0.1, nBands=9, nPhases=10, background_mode='fast', **kwargs):
        self.patDim = None
        self.nTheta = nTheta
        self.nRho = nRho
        self.dTheta = None
        self.dRho = None
        self.rhoMax = None
        self.radonPlan = Radon()
        self.rdnNorm = None
        self.tSigma = tSigma
        self.rSigma = rSigma
        self.kernel = None
        self.peakPad = np.array([11, 11])
        self.padding = np.array([11, 11])
        self.rhoMaskFrac = rhoMaskFrac
        self.rhomask1thresh = None
        self.rhomask1 = None
        self.nBands = nBands
        self.nPhases = 10 if nPhases is None else nPhases
        self.EDAXIQ = False
        self.backgroundsub = None
        self.background_mode = background_mode
        self.useCPU = False
        self._device_plans = {}
        self.dataType = np.dtype([('id', np.int32), ('max', np.float32), ('normmax', np.float32),
                                  ('maxloc', np.float32, (2)), ('avemax', np.float32),
                                  ('aveloc', np.float32, (2)), ('pqmax', np.float32),
                                  ('width', np.float32), ('theta', np.float32), ('rho', np.float32),
                                  ('valid', np.int8), ('band_match_index', np.int64, (self.nPhases,)),
                                  ('bandfit', np.float64, (self.nPhases,))])
        if patterns is not None or patDim is not None:
            self.patDim = np.asarray(patterns.shape[-2:]) if patterns is not None else np.asarray(patDim)
            if 'patternmask' in kwargs:
                self.radonPlan.mask = kwargs.get('patternmask')
            if 'patternmaskindex' in kwargs:
                self.radonPlan.maskindex = kwargs.get('patternmaskindex')
            kw = {k: v for k, v in kwargs.items() if k not in ('patternmask', 'patternmaskindex')}
            self.band_detect_setup(patterns, self.patDim, self.nTheta, self.nRho, self.tSigma,
                                   self.rSigma, self.rhoMaskFrac, self.nBands,
                                   patternmask=self.radonPlan.mask,
                                   patternmaskindex=self.radonPlan.maskindex, **kw)

    # -- set-up (band_detect.py:138-294) ---------------------------------------------------
    def band_detect_setup(self, patterns=None, patDim=None, nTheta=None, nRho=None, tSigma=None,
                          rSigma=None, rhoMaskFrac=None, nBands=None, patternmask=None,
                          patternmaskindex=None, **kwargs):
        p_dim = None
        new_plan = False
        new_kernel = False
        if patterns is None and patDim is not None:
            p_dim = np.asarray(patDim, dtype=np.int64)
        if patterns is not None:
            p_dim = np.asarray(np.shape(patterns)[-2:])
        if p_dim is not None:
            if self.patDim is None or np.sum(np.abs(np.asarray(self.patDim)[-2:] - p_dim[-2:])) != 0:
                new_plan = True
            self.patDim = p_dim
        if nTheta is not None:
            self.nTheta = nTheta
            new_plan = new_kernel = True
        if self.nTheta is not None:
            self.dTheta = 180.0 / self.nTheta
        if nRho is not None:
            self.nRho = nRho
            new_plan = new_kernel = True
        if self.dRho is None:
            new_plan = True
        if patternmask is not None:
            self.radonPlan.mask = patternmask
            new_plan = True
        if patternmaskindex is not None:
            self.radonPlan.maskindex = patternmaskindex
            new_plan = True

        if new_plan and self.patDim is not None:
            pd = np.asarray(self.patDim)
            if 0 < self.rhoMaskFrac < 1:
                self.rhoMax = 0.5 * np.float32(pd.min())
            else:
                self.rhoMax = 0.5 * np.float32(np.sqrt(np.float32(pd[-2]) ** 2 + np.float32(pd[-1]) ** 2))
            self.rhomask1thresh = np.float32(pd.min()) * 0.1
            self.dRho = self.rhoMax / np.float32(self.nRho)
            self.radonPlan.radon_plan_setup(imageDim=self.patDim, nTheta=self.nTheta, nRho=self.nRho,
                                            rhoMax=self.rhoMax, mask=patternmask,
                                            maskindex=patternmaskindex)
            self._setup_rdnmask()
            self._drop_device_state()

        if tSigma is not None:
            self.tSigma = tSigma
            new_kernel = True
        if rSigma is not None:
            self.rSigma = rSigma
            new_kernel = True
        if rhoMaskFrac is not None:
            self.rhoMaskFrac = rhoMaskFrac
        if self.rhoMaskFrac is not None:
            new_kernel = True
        if self.tSigma is None and self.dTheta is not None:
            self.tSigma = 1.0 / self.dTheta
            new_kernel = True
        if self.rSigma is None and self.dRho is not None:
            self.rSigma = 0.5 / np.float32(self.dRho)
            new_kernel = True
        if new_kernel and self.rSigma is not None and self.tSigma is not None:
            old = self.kernel
            self._setup_kernel()
            if old is None or old.shape != self.kernel.shape or not np.array_equal(old, self.kernel):
                self._drop_device_state()   # the reference recomputes the kernel on every call
        self.padding = np.array([np.max([self.peakPad[0], self.padding[0]]),
                                 np.max([self.peakPad[1], self.padding[1]])])
        if nBands is not None:
            if nBands != self.nBands:
                self._drop_device_state()
            self.nBands = nBands

    def _drop_device_state(self):
        """Forget the device plans and host pipelines built for the previous set-up."""
        self._device_plans = {}
        self.__dict__.pop("_pipelines", None)

    def __getstate__(self):
        """Pickled (Ray's object store pickles the indexer and its band-detection plan,
        _ebsd_index_parallel.py:396) without device plans or pipelines: they hold library
        handles, device memory and pinned host memory of this process, and are rebuilt on
        the first call in the process that unpickles."""
        state = self.__dict__.copy()
        state["_device_plans"] = {}
        state.pop("_pipelines", None)
        return state

    def _line_sums(self, img):
        """Unnormalised line sums of one image over the index plan (host set-up only)."""
        plan = self.radonPlan.indexPlan
        nPx = int(np.prod(self.patDim))
        flat = np.concatenate((np.asarray(img, dtype=np.float64).reshape(-1), [0.0, 0.0]))
        valid = plan < nPx
        s = np.where(valid, flat[np.minimum(plan, nPx)], 0.0).sum(axis=2)
        cnt = valid.sum(axis=2)
        return s, cnt

    def _setup_rdnmask(self):
        # band_detect.py:203-251: Radon of the pattern mask without normalisation, > 0,
        # restricted to the central rho rows.
        rp = self.radonPlan
        back = (np.asarray(rp.mask) > 0).astype(np.float32) if rp.mask is not None \
            else np.ones(self.patDim[-2:], dtype=np.float32)
        s, cnt = self._line_sums(back)
        unnorm = np.where(cnt > 0, s.astype(np.float32), np.float32(-1.0))
        self.rhomask1 = (unnorm >= self.rhomask1thresh).astype(np.ubyte)
        rdnmask = unnorm > 0
        if self.rhoMaskFrac >= (1 - 1.0e-12):
            rdnmask = scipyndim.binary_erosion(rdnmask, structure=np.ones((3, 1)),
                                               iterations=int(self.rhoMaskFrac))
        else:
            rdncmask = rdnmask.copy()
            if self.rhoMaskFrac > 0:
                k = max(1, int(self.nRho * self.rhoMaskFrac))
                rdncmask[0:k, :] = 0
                rdncmask[-k:, :] = 0
            elif self.rhoMaskFrac < 0:
                nx, ny = back.shape[-1], back.shape[-2]
                cx = (np.arange(nx) - nx * 0.5) ** 2
                cy = (np.arange(ny) - ny * 0.5) ** 2
                rr = np.sqrt(cx[None, :] + cy[:, None])
                cm = (rr < self.rhoMax * (1.0 + self.rhoMaskFrac)).astype(np.float32)
                cs, ccnt = self._line_sums(cm)
                rdncmask = np.where(ccnt > 0, cs / np.maximum(ccnt, 1), -1.0) > 0
            rdnmask = rdnmask & (rdncmask > 0)
        self.rdnmask = rdnmask > 0

    def _setup_kernel(self):
        # band_detect.py:277-290
        ksz = np.array([np.max([np.int64(4 * self.rSigma), 5]), np.max([np.int64(4 * self.tSigma), 5])])
        ksz = ksz + ((ksz % 2) == 0)
        k = np.zeros(ksz, dtype=np.float32)
        k[(ksz[0] / 2).astype(int), (ksz[1] / 2).astype(int)] = 1
        k = -1.0 * scipyndim.gaussian_filter(k, [self.rSigma, self.tSigma], order=[2, 0])
        k *= 1.0 / np.max(k)
        self.kernel = k.reshape((1, ksz[0], ksz[1]))
        self.peakPad = np.array(np.around([2 * ksz[0], 2 * ksz[1]]), dtype=np.int64)
        self.peakPad += 1 - np.mod(self.peakPad, 2)

    # -- background (band_detect.py:297-408) -------------------------------------------------
    def collect_background(self, fileobj=None, patsIn=None, nsample=None, method='randomStride',
                           sigma=None):
        back = None
        if patsIn is not None:
            pats = patsIn[None] if patsIn.ndim == 2 else patsIn
            npats = pats.shape[0]
            nsample = npats if nsample is None else nsample
            if npats <= nsample:
                back = np.mean(pats, axis=0)[None]
            elif method.upper() == 'RANDOMSTRIDE':
                pick = np.sort(np.random.choice(npats, size=nsample, replace=False))
                back = np.mean(pats[pick], axis=0)
            elif method.upper() == 'EVENSTRIDE':
                back = np.mean(pats[np.arange(0, npats, int(npats / nsample))], axis=0)
        if back is None and fileobj is not None:
            if fileobj.version is None:
                fileobj.read_header()
            npats = fileobj.nPatterns
            nsample = npats if (nsample is None or npats <= nsample) else nsample
            if method.upper() == 'RANDOMSTRIDE':
                pick = np.sort(np.random.choice(npats, size=nsample, replace=False))
            else:
                pick = np.arange(0, npats, int(npats / nsample))
            acc = None
            for i in pick:
                p = fileobj.read_data(convertToFloat=True, patStartCount=[i, 1], returnArrayOnly=True)[0]
                acc = p.astype(np.float64) if acc is None else acc + p
            back = acc / float(len(pick))
        if back is not None:
            back = self.backsub_fit(np.squeeze(back))
        self.backgroundsub = back

    def backsub_fit(self, back, mask=None):
        """2-D Gaussian on a plane fitted inside a circle of radius 0.49*ny, zero-mean;
        falls back to the mean pattern if the fit does not converge (band_detect.py:350-408)."""
        ny, nx = back.shape[-2], back.shape[-1]
        yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
        x, y = xx.ravel(), yy.ravel()

        def surf(x, y, a, x0, y0, sx, sy, c, d, e):
            return a * np.exp(-((x - x0) ** 2) / (2.0 * sx ** 2) - ((y - y0) ** 2) / (2.0 * sy ** 2)) \
                + c + d * x + e * y

        if mask is None:
            r = np.sqrt((np.arange(nx)[None, :] - nx * 0.5) ** 2 + (np.arange(ny)[:, None] - ny * 0.5) ** 2)
            cmask = r < (ny * 0.49)
        else:
            cmask = mask
        sel = np.nonzero(cmask.ravel())[0]
        z = scipyndim.median_filter(np.squeeze(back), 3).ravel()[sel]
        peak = np.unravel_index(back.argmax(), back.shape)
        p0 = [(z.max() - z.min()) * 0.1, peak[1], peak[0], nx / 2.355, ny / 2.355, z.min(), 0, 0]
        try:
            popt, _ = scipyopt.curve_fit(lambda M, *a: surf(M[0], M[1], *a),
                                         np.vstack((x[sel], y[sel])), z, p0)
            fit = surf(x, y, *popt).reshape(ny, nx)
        except RuntimeError:
            print('Warning: no convergence on background gaussian fit ... using mean of the patterns.')
            fit = np.asarray(back, dtype=np.float64).copy()
        fit = fit - np.mean(fit)
        return fit

    # -- device plan ------------------------------------------------------------------------
    def device_plan(self, gpu_id=None):
        import torch
        if not torch.cuda.is_available():
            raise _hip.HipLibraryError("no HIP device visible: BandDetect.find_bands needs a GPU")
        idx = torch.cuda.current_device() if gpu_id is None else int(gpu_id)
        if idx not in self._device_plans:
            self._device_plans[idx] = DevicePlan(self, idx)
        return self._device_plans[idx]

    def _prepare(self, plan, pats):
        """Device tensor + dtype flags for the kernels; float64/other dtypes are converted
        (background subtracted in float64 first, as radon_fast.py:258-259 does)."""
        import torch
        code = {torch.uint8: _hip.EBSD_U8, torch.uint16: _hip.EBSD_U16,
                torch.float32: _hip.EBSD_F32}.get(pats.dtype)
        bg = self.backgroundsub
        plan.set_background(bg, self.background_mode == 'exact')
        if code is not None:
            return pats, 0
        x = pats.to(torch.float64)
        if bg is not None:
            x = x - torch.as_tensor(np.asarray(bg, dtype=np.float64), device=pats.device)
        return x.to(torch.float32), _hip.EBSD_FLAG_NOBG

    def find_bands_device(self, pats, gpu_id=None, out=None, work=None, chunk=None, stream=None):
        """find_bands on a device tensor (N,H,W); returns a dict of device tensors.  The
        kernels run on ``stream`` (a torch.cuda.Stream; default: the current stream); a side
        stream first waits for the current stream, and the outputs are ready in its order."""
        plan = self.device_plan(gpu_id)
        pats = pats.unsqueeze(0) if pats.dim() == 2 else pats
        x, flags = self._prepare(plan, pats)
        return plan.find_bands(x, out=out, work=work, chunk=chunk, flags=flags, stream=stream)

    #: host inputs of at least this many patterns take the pipelined path of find_bands
    STREAM_MIN = 512
    #: that path's patterns per batch (when the caller names no chunksize) and device slots
    #: (at 120x120 uint16: 9.6 ms per 16,384 patterns with 2,048 and 3 slots, against 9.8 with
    #: 2 slots, 10.2 with 4,096 and 12.4 with 1,024)
    STREAM_BATCH = 2048
    STREAM_DEPTH = 3
    #: the pipelined path returns its structured array in pinned host memory (the records are
    #: copied there straight from the device); False returns a pageable copy instead, for
    #: callers that keep many results alive
    PINNED_RESULT = True

    def _find_bands_streamed(self, patternsIn, plan, chunksize):
        """find_bands of host patterns (numpy or CPU tensor, uint8/uint16/float32, at least
        STREAM_MIN of them) through a pinned-host pipeline: batches are copied into a pinned
        ring by a thread pool, uploaded on a copy stream and detected on a compute stream
        while the next batch is copied, and the results come back into pinned memory (the
        structured array's own, packed on the device, when dataType has the reference's
        layout).  Returns the structured array, or None when the input does not take this
        path."""
        import torch
        from .streaming import StreamingBandDetect
        if isinstance(patternsIn, torch.Tensor):
            if patternsIn.is_cuda:
                return None
            arr = patternsIn.numpy()
        else:
            arr = np.asarray(patternsIn)
        if arr.ndim != 3 or arr.shape[0] < self.STREAM_MIN or _dtype_code(arr.dtype) is None:
            return None
        if tuple(arr.shape[1:]) != (plan.H, plan.W):
            raise ValueError(f"patterns must be (N, {plan.H}, {plan.W})")
        arr = np.ascontiguousarray(arr)
        N = int(arr.shape[0])
        batch = int(chunksize) if chunksize and chunksize > 0 else self.STREAM_BATCH
        batch = max(1, min(batch, N))
        # the bands come back as records in the byte layout of dataType when it has the
        # reference's layout: the device writes the returned array itself
        nph = record_phases(self.dataType)
        key = (id(plan), str(arr.dtype), batch, nph, self.STREAM_DEPTH)
        pipes = self.__dict__.setdefault("_pipelines", {})
        sb = pipes.get(key)
        if sb is None:
            pipes.clear()   # one pipeline at a time: its device slots and pinned ring are large
            sb = StreamingBandDetect(self, batch=batch, dtype=str(arr.dtype), gpu_id=plan.device.index,
                                     depth=self.STREAM_DEPTH, chunk=batch, records=nph)
            pipes[key] = sb
        blocks = sb.__dict__.get("_blocks")
        nb = -(-N // batch)
        if blocks is None or len(blocks) < nb:
            blocks = sb._blocks = sb.pinned_blocks(nb)
        if nph is not None:
            # the structured array lives in pinned host memory that the device writes its
            # records into (torch's caching host allocator recycles it once the array is
            # dropped): no host-side assembly and no host copy of the result
            rec = torch.empty(N * sb.record_row_bytes, dtype=torch.uint8, pin_memory=True)
            res = sb.run_array(arr, blocks=blocks[:nb], rec_out=rec)
            bandData = rec.numpy().view(self.dataType).reshape(N, self.nBands)
            if not self.PINNED_RESULT:
                bandData = bandData.copy()
                del rec
        else:
            res = sb.run_array(arr, blocks=blocks[:nb])
            bandData = np.zeros((N, self.nBands), dtype=self.dataType)
        stats = np.empty((N, 3), dtype=np.float32)
        off = 0
        for r in res:
            n = int(r["stats"].shape[0])
            if nph is None:
                for k in ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid"):
                    bandData[k][off:off + n] = r[k].numpy()
            stats[off:off + n] = r["stats"].numpy()
            off += n
        self.last_stats = stats
        return bandData

    def find_bands(self, patternsIn, verbose=0, chunksize=-1, clparams=None, gpu_id=None, **kwargs):
        """band_detect.py:411-526 on the GPU.  ``chunksize`` > 0 bounds the patterns per
        device pass (default 2,048: the Radon of a pass stays in the MI355X's 256 MB
        Infinity Cache at 120x120); ``clparams`` may be a ``HipParam`` whose ``gpu_id``
        selects the device."""
        import torch
        tic0 = timer()
        if gpu_id is None and clparams is not None and hasattr(clparams, "gpu_id"):
            gpu_id = clparams.gpu_id
        plan = self.device_plan(gpu_id)
        fast = self._find_bands_streamed(patternsIn, plan, chunksize)
        if fast is not None:
            if verbose > 0:
                print('Total Band Find Time:', timer() - tic0)
            return fast
        if isinstance(patternsIn, torch.Tensor):
            pats = patternsIn.to(plan.device)
        else:
            arr = np.asarray(patternsIn)
            if arr.ndim == 2:
                arr = arr[None]
            if _dtype_code(arr.dtype) is None and arr.dtype != np.float64:
                arr = arr.astype(np.float64)
            pats = torch.from_numpy(np.ascontiguousarray(arr)).to(plan.device)
        if pats.dim() == 2:
            pats = pats.unsqueeze(0)
        N = pats.shape[0]
        chunk = chunksize if chunksize and chunksize > 0 else None
        with torch.cuda.device(plan.device):
            x, flags = self._prepare(plan, pats)
            out = plan.find_bands(x, chunk=chunk, flags=flags)
            host = {k: v.cpu().numpy() for k, v in out.items()}
        # band_detect.py:425-426 fills band_match_index with -100, but :466 then assigns
        # whole records of the chunk array (np.zeros, :668), so the reference returns 0
        bandData = np.zeros((N, self.nBands), dtype=self.dataType)
        for k in ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid"):
            bandData[k] = host[k]
        if verbose > 0:
            print('Total Band Find Time:', timer() - tic0)
        self.last_stats = host["stats"]
        return bandData
