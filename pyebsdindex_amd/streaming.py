"""Scan-scale streaming: patterns stay in pinned host memory and move to the GPU in
fixed-size batches on a side stream, overlapped with band detection of the previous batch.

The reference indexes a scan by reading chunks of the pattern file on the host and handing
each chunk to ``find_bands`` (pyebsdindex/_ebsd_index_single.py:692-707, chunk loop of
``index_pats``); the OpenCL backend then uploads the chunk and blocks on the result
(opencl/band_detect_cl.py:95-199).  Here each batch is copied with an asynchronous
host-to-device copy on a copy-in stream into one of ``depth`` device slots while the
compute stream runs ``ebsd_find_bands`` on the slot filled before; band results return to
pinned host memory on a separate copy-out stream, so that the next batch's upload never
queues behind the previous batch's results.  Events order the three: a slot is refilled
only after the band detection that read it, and its result buffers are overwritten only
after their copy-back.  All streams keep the default priority.
"""
import numpy as np


class StreamingBandDetect:
    """Pipelined band detection of host-resident pattern batches on one GPU.

    ``bd`` is a ``pyebsdindex_amd.BandDetect`` set up for the pattern size (background and
    mask included); ``batch`` patterns move per copy; ``depth`` device slots (>= 2) let a
    copy overlap the previous batch's band detection."""

    FIELDS = ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid", "stats")

    def __init__(self, bd, batch=8192, dtype="uint16", gpu_id=None, depth=2, chunk=2048):
        import torch
        if depth < 2:
            raise ValueError("depth must be >= 2 (one slot copying while another computes)")
        self.bd = bd
        self.plan = bd.device_plan(gpu_id)
        self.device = self.plan.device
        self.batch = int(batch)
        self.depth = int(depth)
        self.chunk = int(chunk)
        self.tdtype = {"uint8": torch.uint8, "uint16": torch.uint16, "float32": torch.float32}[str(np.dtype(dtype))]
        H, W = self.plan.H, self.plan.W
        with torch.cuda.device(self.device):
            self.slots = [torch.empty((self.batch, H, W), dtype=self.tdtype, device=self.device)
                          for _ in range(self.depth)]
            self.outs = [self.plan.alloc_out(self.batch) for _ in range(self.depth)]
            self.work = self.plan.workspace(self.batch, self.chunk)
            self.copy_stream = torch.cuda.Stream(self.device)      # host -> device batches
            self.compute_stream = torch.cuda.Stream(self.device)
            self.return_stream = torch.cuda.Stream(self.device)    # device -> host results
            self.copied = [torch.cuda.Event() for _ in range(self.depth)]
            self.computed = [torch.cuda.Event() for _ in range(self.depth)]
            self.returned = [torch.cuda.Event() for _ in range(self.depth)]
        self._used = [False] * self.depth

    def pinned_batch(self, n=None):
        """A pinned host tensor shaped for one batch (``n`` patterns, default ``batch``)."""
        import torch
        n = self.batch if n is None else int(n)
        return torch.empty((n, self.plan.H, self.plan.W), dtype=self.tdtype, pin_memory=True)

    def pinned_results(self, n):
        """Pinned host tensors for ``n`` patterns' band results (keys of ``FIELDS``)."""
        import torch
        nB = self.plan.nB
        f = dict(dtype=torch.float32, pin_memory=True)
        shapes = {"maxloc": (n, nB, 2), "aveloc": (n, nB, 2), "stats": (n, 3)}
        res = {}
        for k in self.FIELDS:
            shp = shapes.get(k, (n, nB))
            res[k] = torch.empty(shp, dtype=torch.int8 if k == "valid" else torch.float32, pin_memory=True)
        return res

    def run(self, batches, results=None):
        """Stream ``batches`` (an iterable of pinned host tensors of at most ``batch``
        patterns) through band detection.  Returns pinned host result tensors holding every
        pattern in order (``results`` may pass them in, e.g. from ``pinned_results``)."""
        import torch
        batches = list(batches)
        total = sum(int(b.shape[0]) for b in batches)
        if results is None:
            results = self.pinned_results(total)
        x_flags = None
        off = 0
        with torch.cuda.device(self.device):
            for j, hb in enumerate(batches):
                n = int(hb.shape[0])
                if n > self.batch or tuple(hb.shape[1:]) != (self.plan.H, self.plan.W):
                    raise ValueError("batch larger than the pipeline's batch size or of another pattern size")
                if hb.dtype != self.tdtype:
                    raise TypeError(f"batch dtype {hb.dtype}, pipeline dtype {self.tdtype}")
                s = j % self.depth
                cs, ks = self.copy_stream, self.compute_stream
                if self._used[s]:
                    cs.wait_event(self.computed[s])          # slot s's last band detection has read it
                with torch.cuda.stream(cs):
                    self.slots[s][:n].copy_(hb, non_blocking=True)
                self.copied[s].record(cs)
                ks.wait_event(self.copied[s])
                if self._used[s]:
                    ks.wait_event(self.returned[s])          # slot s's last results have been copied back
                out = self.outs[s]
                sub = {k: v[:n] for k, v in out.items()}
                self.bd.find_bands_device(self.slots[s][:n], gpu_id=self.device.index, out=sub, work=self.work,
                                          chunk=self.chunk, stream=ks)
                self.computed[s].record(ks)
                rs = self.return_stream
                rs.wait_event(self.computed[s])
                with torch.cuda.stream(rs):
                    for k in self.FIELDS:
                        results[k][off:off + n].copy_(sub[k], non_blocking=True)
                self.returned[s].record(rs)
                self._used[s] = True
                off += n
            self.return_stream.synchronize()
            self.copy_stream.synchronize()
        return results
