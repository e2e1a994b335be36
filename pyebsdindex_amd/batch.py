"""Mixed-size batches: many pattern sets, each with its own BandDetect plan, in persistent
Radon and peak launches whose work lists mix the sets (optionally one hipGraph replay).

The reference handles one pattern size per plan and re-plans when the size changes
(pyebsdindex/_ebsd_index_single.py:696-700); here the per-size plans stay resident and one
launch per LDS-occupancy class (Radon) or per Radon geometry (peaks) walks a work list that
mixes them (``ebsd_batch_*`` in include/ebsd_radon.h).  A set of 1,024 patterns or more,
which fills the GPU by itself, runs its own per-size gather and peak launches inside the
same launch sequence and graph.
"""
import ctypes

import numpy as np

from . import _hip


class BatchBandDetect:
    """Bind device pattern tensors of several sizes/dtypes to one batched launch.

    ``detectors``: one ``BandDetect`` per set (its plan, kernel, masks, background);
    ``patterns``: matching device tensors (N_i, H_i, W_i) of uint8/uint16/float32.
    The outputs are per-set dicts of device tensors, as ``BandDetect.find_bands_device``."""

    def __init__(self, detectors, patterns, gpu_id=None):
        import torch
        if len(detectors) != len(patterns) or not detectors:
            raise ValueError("one pattern tensor per detector is required")
        self.lib = _hip.load()
        self.detectors = list(detectors)
        self.plans = [bd.device_plan(gpu_id) for bd in self.detectors]
        self.device = self.plans[0].device
        n = len(self.plans)
        pats, codes = [], []
        for bd, plan, x in zip(self.detectors, self.plans, patterns):
            x = x.unsqueeze(0) if x.dim() == 2 else x
            xx, flags = bd._prepare(plan, x)
            xx, code = plan._check_pats(xx)
            if code is None:
                raise TypeError(f"unsupported device pattern dtype {xx.dtype}")
            pats.append(xx)
            codes.append(code | flags)
        self.patterns = pats            # keep the (possibly converted) inputs alive
        self.N = [int(x.shape[0]) for x in pats]
        self.outs = [plan.alloc_out(nq) for plan, nq in zip(self.plans, self.N)]
        P = ctypes.c_void_p
        self._plan_arr = (P * n)(*[p.handle.value for p in self.plans])
        self._pat_arr = (P * n)(*[x.data_ptr() for x in pats])
        self._dt_arr = (ctypes.c_int * n)(*codes)
        self._n_arr = (ctypes.c_int64 * n)(*self.N)
        self._bands = (_hip.EbsdBands * n)(*[plan.bands_struct(o) for plan, o in zip(self.plans, self.outs)])
        nbytes = int(self.lib.ebsd_batch_workspace_bytes(n, self._plan_arr, self._n_arr))
        self.work = torch.empty(max(nbytes, 4) // 4 + 64, dtype=torch.float32, device=self.device)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _hip.check(self.lib.ebsd_batch_create(n, self._plan_arr, self._pat_arr, self._dt_arr, self._n_arr,
                                                  self._bands, self.work.data_ptr(), self.work.numel() * 4,
                                                  ctypes.byref(h)))
        self.handle = h
        self.graph_stream = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                self.lib.ebsd_batch_destroy(h)
            except Exception:
                pass
            self.handle = None

    def run(self, stream=None):
        """Launch the two persistent kernels on ``stream`` (default: current stream)."""
        s = _hip.launch_stream(stream, self.device)
        _hip.check(self.lib.ebsd_batch_run(self.handle, s.cuda_stream))
        return self.outs

    def graph_launch(self, stream=None):
        """Replay the captured graph (captured on the first call) on a non-default stream.
        Without ``stream`` the replay runs on a private stream ordered both ways with the
        current stream: it starts after the work queued there, and work queued there
        afterwards (reading the outputs, to_numpy) waits for it.  With ``stream`` it runs on
        that (non-default) stream after the current stream's queued work; the outputs are
        then ready in ``stream``'s order."""
        import torch
        cur = torch.cuda.current_stream(self.device)
        if stream is None:
            if self.graph_stream is None:
                self.graph_stream = torch.cuda.Stream(self.device)
            s = self.graph_stream
            s.wait_stream(cur)
            _hip.check(self.lib.ebsd_batch_graph_launch(self.handle, s.cuda_stream))
            cur.wait_stream(s)
        else:
            s = _hip.launch_stream(stream, self.device)
            _hip.check(self.lib.ebsd_batch_graph_launch(self.handle, s.cuda_stream))
        return self.outs

    def to_numpy(self):
        """Structured arrays (the reference dataType) per set."""
        res = []
        for bd, out, nq in zip(self.detectors, self.outs, self.N):
            host = {k: v.cpu().numpy() for k, v in out.items()}
            arr = np.zeros((nq, bd.nBands), dtype=bd.dataType)
            for k in ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid"):
                arr[k] = host[k]
            res.append(arr)
        return res
