"""Scan-scale streaming: patterns stay in pinned host memory and move to the GPU in
fixed-size batches on a side stream, overlapped with band detection of the previous batch.

The reference indexes a scan by reading chunks of the pattern file on the host and handing
each chunk to ``find_bands`` (pyebsdindex/_ebsd_index_single.py:692-707, chunk loop of
``index_pats``); the OpenCL backend then uploads the chunk and blocks on the result
(opencl/band_detect_cl.py:95-199).  Here each batch is copied with an asynchronous
host-to-device copy on a copy-in stream into one of ``depth`` device slots while the
compute stream runs ``ebsd_find_bands`` on a slot filled before.  A slot's band results
live in one device block (all fields) that returns to pinned host memory in a single copy
on a copy-out stream.  Events order the streams: a slot is refilled only after the band
detection that read it, and its result block is overwritten only after its copy-back.

The host enqueues uploads ``depth - 1`` batches ahead of band detection.  Streams can share
a hardware queue (HIP spreads a process's streams over a few queues), and a wait in an
in-order queue holds everything behind it: an upload enqueued after a batch's band
detection would then start only when that detection ends.  Enqueued ahead, an upload
only ever waits behind the detection that last used its slot, which is the dependency it
has anyway.  All streams keep the default priority.

``run_file`` feeds the same pipeline from a UP1/UP2 pattern file: each batch is read with
positioned reads straight into one buffer of a small pinned ring, just before its upload is
enqueued, and a ring buffer is read into again only once the upload that used it has
finished.  The reads of a batch run on the host while the device copies and detects the
batches before it.
"""
import threading

import numpy as np

_pools = {}
_pools_lock = threading.Lock()


def _copy_pool(threads):
    """A kept pool of host threads for the copies into pinned memory (numpy releases the
    interpreter lock while it copies)."""
    from concurrent.futures import ThreadPoolExecutor
    with _pools_lock:
        pool = _pools.get(int(threads))
        if pool is None:
            pool = _pools[int(threads)] = ThreadPoolExecutor(max_workers=int(threads))
        return pool


class StreamingBandDetect:
    """Pipelined band detection of host-resident pattern batches on one GPU.

    ``bd`` is a ``pyebsdindex_amd.BandDetect`` set up for the pattern size (background and
    mask included); ``batch`` patterns move per copy; ``depth`` device slots (>= 2) let a
    copy overlap the previous batch's band detection."""

    FIELDS = ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid", "stats")

    def __init__(self, bd, batch=8192, dtype="uint16", gpu_id=None, depth=3, chunk=None, records=None):
        import torch
        if depth < 2:
            raise ValueError("depth must be >= 2 (one slot copying while another computes)")
        # records: a phase count -> each batch's bands come back as packed records in the byte
        # layout of BandDetect.dataType with that many phases ("records", uint8 (n, nB*R)),
        # beside "stats", instead of one tensor per field
        self.records = None if records is None else int(records)
        self.bd = bd
        self.plan = bd.device_plan(gpu_id)
        self.device = self.plan.device
        self.batch = int(batch)
        self.depth = int(depth)
        # patterns per Radon+peak pass inside a batch (None: the plan's automatic size)
        self.chunk = self.plan.auto_chunk(self.batch) if chunk is None else int(chunk)
        self.tdtype = {"uint8": torch.uint8, "uint16": torch.uint16, "float32": torch.float32}[str(np.dtype(dtype))]
        H, W = self.plan.H, self.plan.W
        nB = self.plan.nB
        # field layout of one result block: (name, torch dtype, per-pattern shape), 256-byte aligned
        self.layout = []
        off = 0
        fields = self.FIELDS if self.records is None else ("records", "stats")
        for k in fields:
            if k == "records":
                shp, dt, esz = (nB * (49 + 16 * self.records),), torch.uint8, 1
            else:
                shp = {"maxloc": (nB, 2), "aveloc": (nB, 2), "stats": (3,)}.get(k, (nB,))
                dt, esz = (torch.int8, 1) if k == "valid" else (torch.float32, 4)
            nbytes = self.batch * int(np.prod(shp)) * esz
            self.layout.append((k, dt, shp, off, nbytes))
            off += (nbytes + 255) // 256 * 256
        self.block_bytes = off
        with torch.cuda.device(self.device):
            self.slots = [torch.empty((self.batch, H, W), dtype=self.tdtype, device=self.device)
                          for _ in range(self.depth)]
            self.blocks = [torch.empty(self.block_bytes, dtype=torch.uint8, device=self.device)
                           for _ in range(self.depth)]
            self.outs = [self._views(b, self.batch) for b in self.blocks]
            if self.records is not None:
                # the fields stay on the device (one set per slot); the block holds the records
                # packed from them and the statistics
                for o in self.outs:
                    rec = o.pop("records")
                    o.update({k: v for k, v in self.plan.alloc_out(self.batch).items() if k != "stats"})
                    o["records"] = rec
            self.work = self.plan.workspace(self.batch, self.chunk)
            self.copy_stream = torch.cuda.Stream(self.device)      # host -> device batches
            self.compute_stream = torch.cuda.Stream(self.device)
            self.return_stream = torch.cuda.Stream(self.device)    # device -> host results
            self.copied = [torch.cuda.Event() for _ in range(self.depth)]
            self.computed = [torch.cuda.Event() for _ in range(self.depth)]
            self.returned = [torch.cuda.Event() for _ in range(self.depth)]
        self._used = [False] * self.depth

    def _views(self, block, n):
        """Typed views of the first ``n`` patterns of every field of a result block."""
        out = {}
        for k, dt, shp, off, nbytes in self.layout:
            v = block[off:off + nbytes].view(dt).view((self.batch,) + shp)
            out[k] = v[:n]
        return out

    def pinned_batch(self, n=None):
        """A pinned host tensor shaped for one batch (``n`` patterns, default ``batch``)."""
        import torch
        n = self.batch if n is None else int(n)
        return torch.empty((n, self.plan.H, self.plan.W), dtype=self.tdtype, pin_memory=True)

    def pinned_blocks(self, nbatches):
        """Pinned host result blocks for ``nbatches`` batches."""
        import torch
        return [torch.empty(self.block_bytes, dtype=torch.uint8, pin_memory=True) for _ in range(nbatches)]

    def run(self, batches, blocks=None):
        """Stream ``batches`` (pinned host tensors of at most ``batch`` patterns each)
        through band detection.  Returns, per batch, a dict of pinned host tensors (views of
        one result block per batch; ``blocks`` may pass the blocks in, e.g. from
        ``pinned_blocks``, to reuse them)."""
        batches = list(batches)
        for hb in batches:
            self._check_batch(hb)
        return self._run(len(batches), lambda j: batches[j], [int(hb.shape[0]) for hb in batches], blocks)

    def run_file(self, patfile, patStart=0, nPatterns=-1, blocks=None, ring=None, threads=8):
        """Stream patterns [patStart, patStart+nPatterns) of a pattern file (``UPFile`` from
        ``pyebsdindex_amd.ebsd_pattern``; -1 = to the end) through band detection.

        Batches are read from the file straight into a ring of pinned host buffers
        (``ring``, default ``depth + 1`` of them, from ``pinned_batch``) with positioned
        reads on ``threads`` threads; a buffer is read again only after the upload that
        used it has finished, so file reads of one batch overlap the copy and band
        detection of the ones before.  Returns the per-batch results of ``run``."""
        import torch
        if np.dtype(patfile.filedatatype) != np.dtype(str(self.tdtype).replace("torch.", "")):
            raise TypeError(f"file holds {np.dtype(patfile.filedatatype)}, pipeline dtype {self.tdtype}")
        if (int(patfile.patternH), int(patfile.patternW)) != (self.plan.H, self.plan.W):
            raise ValueError("file patterns are of another size than the pipeline's")
        total = int(patfile.nPatterns) - int(patStart) if int(nPatterns) < 0 else int(nPatterns)
        if patStart < 0 or total < 0 or patStart + total > int(patfile.nPatterns):
            raise ValueError("pattern range outside the file")
        counts = [min(self.batch, total - o) for o in range(0, total, self.batch)]
        if ring is None:
            ring = [self.pinned_batch() for _ in range(self.depth + 1)]
        if len(ring) < 2:
            raise ValueError("the host ring needs at least two buffers")
        for hb in ring:
            if int(hb.shape[0]) < self.batch or not hb.is_pinned():
                raise ValueError("ring buffers must be pinned and hold a whole batch")
        R = len(ring)
        freed = [None] * R                    # event: last upload out of ring buffer r has finished

        def get(j):
            r, n = j % R, counts[j]
            if freed[r] is not None:
                freed[r].synchronize()
            patfile.read_into(ring[r], int(patStart) + j * self.batch, n, threads=threads)
            return ring[r][:n]

        def uploaded(j):
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
            freed[j % R] = ev

        return self._run(len(counts), get, counts, blocks, uploaded)

    def run_array(self, arr, blocks=None, ring=None, threads=8, rec_out=None):
        """Stream a host array (N, H, W) of the pipeline's dtype (numpy, or a CPU tensor)
        through band detection: each batch is copied by ``threads`` threads into one buffer
        of a pinned host ring (``ring``, default ``depth + 1`` buffers from
        ``pinned_batch``) just before its upload, and a ring buffer is written again only
        after the upload that read it has finished, so the host copy of one batch overlaps
        the upload and band detection of the ones before.  Returns the results of ``run``.
        ``rec_out`` (pipelines made with ``records``): a pinned uint8 tensor of N records
        rows, ``N * record_row_bytes`` bytes, that receives every batch's records straight
        from the device (the results' own "records" views are then not filled)."""
        import torch
        if isinstance(arr, torch.Tensor):
            arr = arr.numpy()
        arr = np.asarray(arr)
        if arr.ndim != 3 or tuple(arr.shape[1:]) != (self.plan.H, self.plan.W):
            raise ValueError(f"patterns must be (N, {self.plan.H}, {self.plan.W})")
        if arr.dtype != np.dtype(str(self.tdtype).replace("torch.", "")):
            raise TypeError(f"array holds {arr.dtype}, pipeline dtype {self.tdtype}")
        total = int(arr.shape[0])
        counts = [min(self.batch, total - o) for o in range(0, total, self.batch)]
        if ring is None:
            ring = self.host_ring()
        R = len(ring)
        if R < 2:
            raise ValueError("the host ring needs at least two buffers")
        freed = [None] * R
        pool = _copy_pool(threads)

        def get(j):
            r, n = j % R, counts[j]
            if freed[r] is not None:
                freed[r].synchronize()
            dst = ring[r][:n].numpy()
            src = arr[j * self.batch:j * self.batch + n]
            step = max(1, -(-n // max(1, threads)))
            jobs = [pool.submit(np.copyto, dst[o:o + step], src[o:o + step]) for o in range(0, n, step)]
            for f in jobs:
                f.result()
            return ring[r][:n]

        def uploaded(j):
            ev = torch.cuda.Event()
            ev.record(self.copy_stream)
            freed[j % R] = ev

        return self._run(len(counts), get, counts, blocks, uploaded, rec_out)

    def host_ring(self):
        """The pipeline's own pinned host ring (``depth + 1`` batches), made on first use."""
        ring = getattr(self, "_ring", None)
        if ring is None:
            ring = self._ring = [self.pinned_batch() for _ in range(self.depth + 1)]
        return ring

    def _check_batch(self, hb):
        if int(hb.shape[0]) > self.batch or tuple(hb.shape[1:]) != (self.plan.H, self.plan.W):
            raise ValueError("batch larger than the pipeline's batch size or of another pattern size")
        if hb.dtype != self.tdtype:
            raise TypeError(f"batch dtype {hb.dtype}, pipeline dtype {self.tdtype}")

    @property
    def record_row_bytes(self):
        """Bytes of one pattern's records (nB records of 49 + 16 * phases bytes)."""
        return None if self.records is None else self.plan.nB * (49 + 16 * self.records)

    def _run(self, nb, get_batch, counts, blocks, on_upload=None, rec_out=None):
        """The pipeline over ``nb`` batches: ``get_batch(j)`` returns batch j as a pinned host
        tensor of ``counts[j]`` patterns, called in order just before its upload is
        enqueued; ``on_upload(j)`` runs right after."""
        import torch
        if blocks is None:
            blocks = self.pinned_blocks(nb)
        if len(blocks) < nb:
            raise ValueError("fewer result blocks than batches")
        if rec_out is not None:
            rb = self.record_row_bytes
            if rb is None:
                raise ValueError("rec_out needs a pipeline made with records")
            if (rec_out.dtype != torch.uint8 or rec_out.is_cuda or not rec_out.is_pinned()
                    or not rec_out.is_contiguous() or rec_out.numel() < sum(counts) * rb):
                raise ValueError(f"rec_out must be a pinned contiguous uint8 tensor of {sum(counts) * rb} bytes")
            # the statistics' span of a result block (the records lead it, at offset 0)
            st = next((off, nbytes) for k, dt, shp, off, nbytes in self.layout if k == "stats")
        cs, ks, rs = self.copy_stream, self.compute_stream, self.return_stream
        D = self.depth

        def upload(j):
            s = j % D
            hb = get_batch(j)
            self._check_batch(hb)
            if self._used[s]:
                cs.wait_event(self.computed[s])              # slot s's last band detection has read it
            with torch.cuda.stream(cs):
                self.slots[s][:hb.shape[0]].copy_(hb, non_blocking=True)
            self.copied[s].record(cs)
            if on_upload is not None:
                on_upload(j)

        results = []
        ahead = D - 1
        with torch.cuda.device(self.device):
            for j in range(min(ahead, nb)):
                upload(j)
            for j in range(nb):
                if j + ahead < nb:
                    upload(j + ahead)                        # its slot's last reader is batch j - 1
                s, n = j % D, int(counts[j])
                ks.wait_event(self.copied[s])
                if self._used[s]:
                    ks.wait_event(self.returned[s])          # slot s's last results are back on the host
                sub = {k: v[:n] for k, v in self.outs[s].items() if k != "records"}
                self.bd.find_bands_device(self.slots[s][:n], gpu_id=self.device.index, out=sub, work=self.work,
                                          chunk=self.chunk, stream=ks)
                if self.records is not None:
                    self.plan.pack_records(sub, self.records, dst=self.outs[s]["records"], stream=ks)
                self.computed[s].record(ks)
                self._used[s] = True
                rs.wait_event(self.computed[s])
                with torch.cuda.stream(rs):
                    if rec_out is None:
                        blocks[j].copy_(self.blocks[s], non_blocking=True)
                    else:
                        o0 = j * self.batch * rb
                        rec_out[o0:o0 + n * rb].copy_(self.blocks[s][:n * rb], non_blocking=True)
                        blocks[j][st[0]:st[0] + st[1]].copy_(self.blocks[s][st[0]:st[0] + st[1]], non_blocking=True)
                self.returned[s].record(rs)
                results.append(self._views(blocks[j], n))
            rs.synchronize()
            cs.synchronize()
        return results

    @staticmethod
    def concat(results):
        """Per-field host tensors of all patterns from ``run``'s per-batch results."""
        import torch
        return {k: torch.cat([r[k] for r in results]) for k in results[0]} if results else {}
