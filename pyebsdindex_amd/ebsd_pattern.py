"""Raw pattern files (EDAX .up1/.up2, Oxford .ebsp, ThermoFisher .pat): the reference's
reader interface, plus reads straight into pinned host memory for the streaming pipeline.

Mirrors ``pyebsdindex.ebsd_pattern`` for the raw formats (``get_pattern_file_obj``
ebsd_pattern.py:33-116, ``pat_flt2int`` :118-181, ``EBSDPatterns`` :184-203,
``EBSDPatternFile`` :208-431, ``UPFile`` :435-708, ``EBSPFile`` :709-1089, ``TFPatFile``
:1090-1396): the same header fields, the same ``read_data`` / ``pat_reader`` semantics and
return values (patterns plus scan x/y locations), the same writers.  The HDF5 families need
h5py, which this environment lacks; they raise NotImplementedError.

File layouts, little-endian, as the reference reads them:

* UP, version 1: uint32 version, W, H, data offset; patterns follow.  The pattern count is
  (file size - 16) / (W * H * bytes per pixel); the scan is one row unless ``nCols`` is set.
* UP, version >= 3: uint32 version, W, H, data offset; uint8 extra patterns; uint32 nCols,
  nRows; uint8 hex flag; float64 x step, y step (42 bytes); patterns at the offset.  Pixels
  are uint8 (.up1) or uint16 (.up2), row-major, patterns in scan order.
* EBSP: int64 holding -version (none for version 0); from version 4 one more byte; then a
  table of uint64 file offsets, one per pattern (0 = pattern not stored).  At each offset:
  a header of 4 uint32 (6 from version 5) whose last three are H, W and the pattern's byte
  count (the first is 1 for a compressed pattern), the pixels, then -- when the file holds
  positions -- float64 x and float64 y, each preceded by a flag byte from version 2 on (a
  version 2+ file without positions has one 0 byte there).  The pattern count is the
  number of table entries before the first pattern's data (version <= 4) or the table slot
  that holds H + (W << 32) (version 5).  The scan's columns and steps come from the
  stored positions.
* TF .pat: uint32 version, nCols, nRows, W, H; int16 scan flag (< 0: a full scan of
  nCols * nRows patterns, else the count comes from the file size); uint8 bit depth; int16
  background flag (< 0: one uint16 background pattern follows the 25-byte header).  Then one
  record per pattern: uint32 column, uint32 row, uint16 pixels.  A ``.ebsd`` file beside
  it holds the step size as float64 45 bytes before its end.

``read_into`` copies a run of patterns from the file into a caller's buffer (a pinned host
tensor of ``StreamingBandDetect``) with positioned vector reads on several threads: the
pixel bytes go from the page cache into the buffer once, with no intermediate array; the
per-pattern headers and footers between them are read into a small scratch buffer.
"""
import os
import shutil
import struct
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

try:
    _IOV_MAX = int(os.sysconf("SC_IOV_MAX"))
except (ValueError, OSError, AttributeError):
    _IOV_MAX = 1024
_IOV_MAX = max(16, min(_IOV_MAX, 1024))


def get_pattern_file_obj(path, file_type=str('')):
    """Pattern-file object for ``path``, its type from ``file_type`` or the extension, with
    its header read (ebsd_pattern.py:33-116).  Raises ValueError for an unknown extension
    and NotImplementedError for the HDF5 formats, which need h5py."""
    filepath = np.atleast_1d(path)[0]
    ftype = file_type
    if ftype == str(''):
        ext = str.lower(Path(filepath).suffix)
        ftype = {'.up1': 'UP1', '.up2': 'UP2', '.pat': 'TFPAT', '.ebsp': 'EBSP', '.oh5': 'OH5',
                 '.h5': 'H5', '.h5oina': 'H5OINA', '.dm5': 'DM5'}.get(ext)
        if ftype is None:
            raise ValueError('Error: extension not recognized')
    kind = ftype.upper()
    if kind in ('UP1', 'UP2'):
        obj = UPFile(path)
    elif kind == 'EBSP':
        obj = EBSPFile(path)
    elif kind == 'TFPAT':
        obj = TFPatFile(path)
    else:
        raise NotImplementedError(f"pattern file type {ftype}: the HDF5 formats need h5py, which is not available")
    if obj.read_header() != 0:
        raise OSError(f"cannot read the header of {filepath}")
    return obj


def pat_flt2int(patterns, typeout=None, method='clip', scalevalue=0.98, maxScale=None):
    """Patterns (an array or ``EBSDPatterns``) as unsigned integers (ebsd_pattern.py:118-181).

    ``typeout`` None picks uint8, or uint16 when the largest value exceeds 258; a floating
    ``typeout`` only converts.  ``method``: 'clip' clips to [0, type max] (or ``maxScale``
    if smaller); 'fullscale' maps [min, min + maxScale] onto [0, scalevalue * type max],
    rounded; 'scale' first lifts a negative minimum to 0, then multiplies by
    scalevalue * type max / maxScale and rounds.  ``maxScale`` defaults to the largest
    (shifted) value.  Returns (n, H, W)."""
    pats = patterns.patterns if isinstance(patterns, EBSDPatterns) else np.asarray(patterns)
    if pats.ndim == 2:
        pats = pats[None]
    hi = pats.max()
    lo = pats.min()
    if typeout is None:
        typeout = np.uint16 if hi > 258 else np.uint8
    typeout = np.dtype(typeout).type
    if np.issubdtype(typeout, np.floating):
        return pats.astype(typeout)
    top = 65535 if typeout == np.uint16 else 255
    out = np.zeros(pats.shape, dtype=typeout)
    if method == 'clip':
        cap = top if maxScale is None else min(maxScale, top)
        out[...] = pats.clip(0, cap).astype(typeout)
    elif method == 'fullscale':
        t = pats.astype(np.float32) - lo
        scale = t.max() if maxScale is None else maxScale
        t *= scalevalue * top / scale
        out[...] = np.around(t.clip(0, top)).astype(typeout)
    elif method == 'scale':
        t = pats.astype(np.float32)
        if lo < 0:
            t += 0 - lo
        scale = t.max() if maxScale is None else maxScale
        t *= scalevalue * top / scale
        out[...] = np.around(t).astype(typeout)
    return out


class EBSDPatterns:
    """A block of patterns read from a file with their scan locations (ebsd_pattern.py:184-203)."""

    def __init__(self, path=None):
        self.vendor = None
        self.filepath = None
        self.filetype = None
        self.patternW = None
        self.patternH = None
        self.nFileCols = None
        self.nFileRows = None
        self.nPatterns = None
        self.hexflag = None
        self.xStep = None
        self.yStep = None
        self.patStart = [0, 0]
        self.patterns = None
        self.xyLocations = None


def _preadv_all(fd, bufs, offset, path):
    """Fill the buffers in order from ``offset`` on, however the kernel splits the reads."""
    k, n = 0, len(bufs)
    while k < n:
        got = os.preadv(fd, bufs[k:k + _IOV_MAX], offset)
        if got <= 0:
            raise OSError(f"short read at byte {offset} of {path}")
        offset += got
        while k < n and got >= len(bufs[k]):
            got -= len(bufs[k])
            k += 1
        if got:
            bufs[k] = bufs[k][got:]


class EBSDPatternFile:
    """Base of the pattern-file classes (ebsd_pattern.py:208-431)."""

    def __init__(self, path, filetype=None):
        self.filepath = None
        self.set_filepath(path)
        self.vendor = None
        self.version = None
        self.nCols = None
        self.nRows = None
        self.nPatterns = None
        self.patternW = None
        self.patternH = None
        self.xStep = None
        self.yStep = None
        self.xyCenter = np.array([0.0, 0.0])
        self.datamax = None
        self.datamin = None
        self.hexflag = False
        self.filetype = filetype
        self.filedatatype = np.uint8

    def set_filepath(self, path=None):
        if path is not None:
            self.filepath = Path(np.atleast_1d(path)[0]).expanduser().resolve()

    def read_header(self, path=None):
        pass

    def write_header(self, writeBlank=False, bitdepth=None):
        pass

    # -- reading -----------------------------------------------------------------------
    @property
    def pattern_bytes(self):
        return int(self.patternW) * int(self.patternH) * np.dtype(self.filedatatype).itemsize

    def _runs(self, patStart, n):
        """Where the pixels of patterns [patStart, patStart+n) lie: a list of (file offset of
        a run's first pattern, bytes from one pattern's pixels to the next's, index of the
        run's first pattern in the request, patterns in the run).  Patterns the file does
        not store are in no run."""
        raise NotImplementedError

    def _xy(self, patStart, n):
        """Scan locations of patterns [patStart, patStart+n), as ``pat_reader`` returns them."""
        raise NotImplementedError

    def _pool(self, n):
        """A reader thread pool of at least ``n`` threads, kept between calls."""
        pool = self.__dict__.get("_readers")
        if pool is None or pool._max_workers < n:
            if pool is not None:
                pool.shutdown(wait=True)
            pool = self._readers = ThreadPoolExecutor(n)
        return pool

    def __getstate__(self):
        state = self.__dict__.copy()
        state.pop("_readers", None)
        return state

    def read_into(self, dst, patStart, nPatToRead, threads=8):
        """Copy patterns [patStart, patStart+n) from the file into ``dst`` (a C-contiguous
        numpy array or CPU torch tensor, pinned or not, of at least n patterns of the file's
        size and type), with positioned reads on up to ``threads`` threads; patterns the
        file does not store are zero.  Raises ValueError for a short buffer, a wrong type
        or a range past the end of the file."""
        n = int(nPatToRead)
        arr = dst.numpy() if hasattr(dst, 'numpy') else dst
        if not arr.flags['C_CONTIGUOUS']:
            raise ValueError("destination must be contiguous")
        if arr.dtype != np.dtype(self.filedatatype):
            raise ValueError(f"destination dtype {arr.dtype}, file holds {np.dtype(self.filedatatype)}")
        pb = self.pattern_bytes
        if arr.nbytes < n * pb:
            raise ValueError("destination holds fewer than the requested patterns")
        if patStart < 0 or patStart + n > int(self.nPatterns):
            raise ValueError(f"patterns [{patStart}, {patStart + n}) outside the file's {int(self.nPatterns)}")
        if n == 0:
            return dst
        view = memoryview(arr.reshape(-1).view(np.uint8))[:n * pb]
        runs = self._runs(int(patStart), n)
        covered = sum(r[3] for r in runs)
        if covered < n:                          # patterns not stored read as zeros
            stored = np.zeros(n, bool)
            for _, _, i0, cnt in runs:
                stored[i0:i0 + cnt] = True
            flat = arr.reshape(-1).view(np.uint8)[:n * pb].reshape(n, pb)
            flat[~stored] = 0
        # jobs: (file offset, list of buffers) -- contiguous runs in pieces of >= 4 MiB,
        # strided runs in pieces of at most IOV_MAX / 2 patterns (pixels, gap, pixels, ...)
        total = covered * pb
        nthr = max(1, min(int(threads), total // (4 << 20) or 1))
        jobs = []
        for off, stride, i0, cnt in runs:
            if stride == pb or cnt == 1:
                nbytes = cnt * pb
                step = max(4 << 20, -(-nbytes // nthr))
                for o in range(0, nbytes, step):
                    e = min(nbytes, o + step)
                    jobs.append((off + o, [view[i0 * pb + o:i0 * pb + e]]))
            else:
                gap = stride - pb
                per = max(1, min(_IOV_MAX // 2, -(-cnt // nthr)))
                for k0 in range(0, cnt, per):
                    k1 = min(cnt, k0 + per)
                    junk = memoryview(bytearray(gap))
                    bufs = []
                    for k in range(k0, k1):
                        bufs.append(view[(i0 + k) * pb:(i0 + k + 1) * pb])
                        if k + 1 < k1:
                            bufs.append(junk)
                    jobs.append((off + k0 * stride, bufs))
        fd = os.open(Path(self.filepath).expanduser(), os.O_RDONLY)
        try:
            path = self.filepath
            if nthr == 1 or len(jobs) == 1:
                for o, bufs in jobs:
                    _preadv_all(fd, bufs, o, path)
            else:
                list(self._pool(nthr).map(lambda j: _preadv_all(fd, j[1], j[0], path), jobs))
        finally:
            os.close(fd)
        return dst

    def pat_reader(self, patStart=0, nPatToRead=1):
        """(patterns (n, H, W) of the file's type, scan locations (n, 2)) for patterns
        [patStart, patStart+n)."""
        out = np.empty((int(nPatToRead), int(self.patternH), int(self.patternW)), dtype=self.filedatatype)
        self.read_into(out, patStart, nPatToRead)
        return out, self._xy(int(patStart), int(nPatToRead))

    def read_data(self, path=None, convertToFloat=False, patStartCount=[0, -1], returnArrayOnly=False):
        """Read a run of patterns ([start, count], count -1 = to the end, 0 = one) or a slab
        ([[col0, row0], [ncols, nrows]], negative = to the edge); counts are clipped to the
        scan (ebsd_pattern.py:236-329).  Returns (patterns, xy) with ``returnArrayOnly``,
        else an ``EBSDPatterns``."""
        if path is not None:
            self.set_filepath(path)
            self.read_header()
        if self.version is None:
            self.read_header()
        typeout = np.float32 if convertToFloat else self.filedatatype
        se = np.asarray(patStartCount, dtype=np.int64)
        if se.ndim == 1:
            patStart = int(se[0])
            nread = int(se[-1])
            if nread == -1:
                nread = int(self.nPatterns) - patStart
            if nread == 0:
                nread = 1
            if patStart + nread > int(self.nPatterns):
                nread = int(self.nPatterns) - patStart
            pats, xyloc = self.pat_reader(patStart, nread)
            patterns = pats.astype(typeout)
            nPatToRead = nread
        elif se.ndim == 2:
            col0, ncol = int(se[0, 0]), int(se[1, 0])
            row0, nrow = int(se[0, 1]), int(se[1, 1])
            patStart = [col0, row0]
            if ncol < 0:
                ncol = int(self.nCols) - col0
            if nrow < 0:
                nrow = int(self.nRows) - row0
            ncol = min(ncol, int(self.nCols) - col0)
            nrow = min(nrow, int(self.nRows) - row0)
            nPatToRead = [np.uint64(ncol), np.uint64(nrow)]
            patterns = np.zeros((ncol * nrow, int(self.patternH), int(self.patternW)), dtype=typeout)
            xyloc = None
            for i in range(nrow):
                p0 = (row0 + i) * int(self.nCols) + col0
                pt, xy = self.read_data(convertToFloat=convertToFloat, patStartCount=[p0, ncol],
                                        returnArrayOnly=True)
                if xyloc is None:
                    xyloc = np.zeros((ncol * nrow, 2), dtype=xy.dtype)
                patterns[i * ncol:(i + 1) * ncol] = pt
                xyloc[i * ncol:(i + 1) * ncol] = xy
            if xyloc is None:
                xyloc = np.zeros((0, 2), dtype=np.float32)
        else:
            raise ValueError("patStartCount must be [start, count] or [[col, row], [ncols, nrows]]")
        if returnArrayOnly:
            return patterns, xyloc
        out = EBSDPatterns()
        out.vendor = self.vendor
        out.file = Path(self.filepath).expanduser()
        out.filetype = self.filetype
        out.patternW = self.patternW
        out.patternH = self.patternH
        out.nFileCols = np.uint64(self.nCols)
        out.nFileRows = np.uint64(self.nRows)
        out.nPatterns = np.array(nPatToRead)
        out.hexflag = self.hexflag
        out.xStep = self.xStep
        out.yStep = self.yStep
        out.patStart = np.array(patStart)
        out.patterns = patterns
        out.xyLocations = xyloc
        return out

    def _xyloc(self, patStart, n):
        """Scan-grid locations of patterns [patStart, patStart+n) relative to the scan
        centre (ebsd_pattern.py:545-553)."""
        idx = np.arange(np.int64(patStart), np.int64(patStart + n), dtype=np.uint64)
        yx = np.unravel_index(idx, (np.int64(self.nRows), np.int64(self.nCols)))
        xy = np.array([yx[1], yx[0]]).T.copy().astype(np.float32)
        xy[:, 0] -= self.nCols * 0.5
        xy[:, 1] -= self.nRows * 0.5
        xy[:, 0] *= self.xStep
        xy[:, 1] *= self.yStep
        return xy

    # -- writing -----------------------------------------------------------------------
    def pat_writer(self, pat2write, patStart, nPatToWrite, typewrite):
        pass

    def write_data(self, newpatterns=None, patStartCount=[0, -1], writeHead=False, flt2int='clip',
                   scalevalue=0.98, maxScale=None):
        """Write patterns (an array or ``EBSDPatterns``) as a run ([start, count]) or a slab
        ([[col0, row0], [ncols, nrows]]) of the scan, converted to the file's type by
        ``pat_flt2int`` (ebsd_pattern.py:339-406).  A file that does not exist yet gets its
        header and blank patterns first."""
        writeblank = False
        if not os.path.isfile(Path(self.filepath).expanduser().resolve()):
            writeHead = True
            writeblank = True
        if writeHead:
            self.write_header(writeBlank=writeblank)
        if isinstance(newpatterns, EBSDPatterns):
            pats = newpatterns.patterns
        else:
            pats = np.asarray(newpatterns)
            if pats.ndim == 2:
                pats = pats[None]
        npats = pats.shape[0]
        top = pats.max() if maxScale is None else maxScale
        se = np.asarray(patStartCount)
        if se.ndim == 1:
            patStart = int(se[0])
            nwrite = int(se[-1])
            if nwrite == -1:
                nwrite = npats
            if nwrite == 0:
                nwrite = 1
            if patStart + nwrite > int(self.nPatterns):
                nwrite = int(self.nPatterns) - patStart
            pat2write = pat_flt2int(pats, typeout=self.filedatatype, method=flt2int, scalevalue=scalevalue,
                                    maxScale=None)
            self.pat_writer(pat2write, patStart, nwrite, self.filedatatype)
        elif se.ndim == 2:
            col0, ncol = int(se[0, 0]), int(se[1, 0])
            row0, nrow = int(se[0, 1]), int(se[1, 1])
            if ncol < 0:
                ncol = int(self.nCols) - col0
            if nrow < 0:
                nrow = int(self.nRows) - row0
            ncol = min(ncol, int(self.nCols) - col0)
            nrow = min(nrow, int(self.nRows) - row0)
            for i in range(nrow):
                p0 = (row0 + i) * int(self.nCols) + col0
                self.write_data(newpatterns=pats[i * ncol:(i + 1) * ncol], patStartCount=[p0, ncol],
                                writeHead=False, flt2int=flt2int, scalevalue=0.98, maxScale=top)

    def copy_file(self, newpath, **kwargs):
        """Copy the file to ``newpath`` (ebsd_pattern.py:411-419)."""
        src = Path(self.filepath).expanduser().resolve()
        if newpath is not None:
            dst = Path(np.atleast_1d(newpath)[0]).expanduser().resolve()
        else:
            dst = Path(str(src) + '.copy')
        shutil.copyfile(src, dst)
        return dst

    def _copy_header_only(self, newpath, head_bytes):
        """A copy of the first ``head_bytes`` bytes, zero to the source's size (sparse)."""
        src = Path(self.filepath).expanduser().resolve()
        dst = Path(np.atleast_1d(newpath)[0]).expanduser().resolve() if newpath is not None \
            else Path(str(src) + '.copy')
        with open(src, 'rb') as f:
            head = f.read(int(head_bytes))
            size = f.seek(0, 2)
        with open(dst, 'wb') as f:
            f.write(head)
            f.truncate(size)
        return dst

    def copy_obj(self):
        import copy
        return copy.deepcopy(self)

    def set_scan_rc(self, rc=(0, 0)):
        """Scan rows and columns for files that do not record them (ebsd_pattern.py:421-425)."""
        self.nCols = np.uint64(rc[1])
        self.nRows = np.uint64(rc[0])
        self.nPatterns = np.uint64(self.nCols * self.nRows)

    def write_datamaxmin(self):
        pass

    def read_datamaxmin(self):
        pass


class UPFile(EBSDPatternFile):
    """EDAX .up1 (uint8) / .up2 (uint16) raw pattern file (ebsd_pattern.py:435-708)."""

    def __init__(self, path=None):
        EBSDPatternFile.__init__(self, path)
        self.filetype = 'UP'
        self.vendor = 'EDAX'
        self.filedatatype = None
        self.filePos = None
        self.extraPatterns = 0
        self.hexflag = 0

    def _set_type(self, bitdepth):
        ext = str.lower(Path(self.filepath).suffix)
        if ext == '.up1':
            self.filedatatype = np.uint8
        elif ext == '.up2':
            self.filedatatype = np.uint16
        elif bitdepth == 8:
            self.filedatatype = np.uint8
        elif bitdepth == 16:
            self.filedatatype = np.uint16
        elif self.filedatatype is None:
            raise ValueError('Error: extension not recognized, set "bitdepth" parameter')

    def read_header(self, path=None, bitdepth=None):
        """Header fields as the reference reads them (ebsd_pattern.py:449-513); returns 0,
        or -1 if the file cannot be opened."""
        if path is not None:
            self.set_filepath(path)
        self._set_type(bitdepth)
        fp = Path(self.filepath).expanduser()
        try:
            with open(fp, 'rb') as f:
                head = f.read(42)
        except OSError:
            print("File Not Found:", str(fp))
            return -1
        self.version = np.uint32(struct.unpack_from('<I', head, 0)[0])
        W, H, pos = struct.unpack_from('<3I', head, 4)
        self.patternW, self.patternH, self.filePos = np.uint32(W), np.uint32(H), np.uint32(pos)
        if self.version == 1:
            nbytes = int(W) * int(H) * np.dtype(self.filedatatype).itemsize
            self.nPatterns = np.int64((fp.stat().st_size - 16) // nbytes)
            if self.xStep is None:
                self.xStep = 0.0
            if self.yStep is None:
                self.yStep = 0.0
            if self.nCols is None or self.nCols == 0:
                self.nCols = np.uint64(self.nPatterns)
            if self.nRows is None:
                self.nRows = np.uint64(np.floor(self.nPatterns / self.nCols))
        elif self.version >= 3:
            self.extraPatterns = np.uint8(head[16])
            nc, nr = struct.unpack_from('<2I', head, 17)
            self.nCols, self.nRows = np.uint64(nc), np.uint64(nr)
            self.nPatterns = np.int64(self.nCols * self.nRows)
            self.hexflag = np.uint8(head[25])
            self.xStep, self.yStep = struct.unpack_from('<2d', head, 26)
        return 0

    def _runs(self, patStart, n):
        pb = self.pattern_bytes
        return [(int(self.filePos) + patStart * pb, pb, 0, n)]

    def _xy(self, patStart, n):
        return self._xyloc(patStart, n)

    def write_header(self, writeBlank=False, bitdepth=None):
        """Write the header (version 3 unless set; ebsd_pattern.py:557-627); with
        ``writeBlank`` also zero patterns for the whole scan."""
        self._set_type(bitdepth)
        fp = Path(self.filepath).expanduser()
        if self.version is None:
            self.version = 3
        if self.version == 1:
            if self.filePos is None:
                self.filePos = 16
            head = struct.pack('<4I', int(self.version), int(self.patternW), int(self.patternH), int(self.filePos))
        else:
            if self.filePos is None:
                self.filePos = 42
            head = struct.pack('<4IB2IB2d', int(self.version), int(self.patternW), int(self.patternH),
                               int(self.filePos), int(self.extraPatterns), int(self.nCols), int(self.nRows),
                               int(self.hexflag), float(self.xStep), float(self.yStep))
        with open(fp, 'r+b' if fp.is_file() else 'w+b') as f:
            f.seek(0)
            f.write(head)
            if writeBlank:                  # zero patterns for the whole scan
                f.truncate(int(self.filePos))
                f.truncate(int(self.filePos) + int(self.nRows) * int(self.nCols) * self.pattern_bytes)

    def pat_writer(self, pat2write, patStart, nPatToWrite, typewrite):
        """Write ``nPatToWrite`` patterns at pattern index ``patStart`` (ebsd_pattern.py:629-646)."""
        fp = Path(self.filepath).expanduser()
        nbytes = int(self.patternW) * int(self.patternH) * np.dtype(typewrite).itemsize
        with open(fp, 'r+b') as f:
            f.seek(nbytes * int(patStart) + int(self.filePos))
            np.ascontiguousarray(pat2write[0:nPatToWrite], dtype=typewrite).tofile(f)

    def copy_file(self, newpath, **kwargs):
        """Copy the file; with ``empty_data=True`` only its header, the patterns zero
        (ebsd_pattern.py:684-708)."""
        if kwargs.get('empty_data'):
            return self._copy_header_only(newpath, int(self.filePos))
        return EBSDPatternFile.copy_file(self, newpath)


class EBSPFile(EBSDPatternFile):
    """Oxford .ebsp raw pattern file (ebsd_pattern.py:709-1089)."""

    def __init__(self, path=None):
        EBSDPatternFile.__init__(self, path)
        self.filetype = 'EBSP'
        self.vendor = 'OXFORD'
        self.filedatatype = None
        self.filePos = None           # file offset of each pattern's record (0: not stored)
        self.hasxypos = False
        self._xpos = None
        self._ypos = None

    def _table_start(self):
        if self.version == 0:
            return 0
        return 9 if self.version >= 4 else 8

    def _head_bytes(self):
        return 24 if self.version >= 5 else 16

    def read_header(self, path=None, bitdepth=None):
        """Header, offset table, pattern size and type, and the scan from the stored
        positions (ebsd_pattern.py:726-908).  Returns 0, or -1 if the file cannot be opened;
        raises NotImplementedError for compressed patterns, ValueError for a table that
        does not describe a file of this format."""
        if path is not None:
            self.set_filepath(path)
        fp = Path(self.filepath).expanduser()
        try:
            f = open(fp, 'rb')
        except OSError:
            print("File Not Found:", str(fp))
            return -1
        with f:
            first = np.frombuffer(f.read(8), dtype=np.int64)[0]
            self.version = int(-first) if -first > 0 else 0
            if self.version == 0:
                raise ValueError(f"{fp}: EBSP files without a version word are not read (nor by the reference)")
            t0 = self._table_start()
            size = f.seek(0, 2)
            # the first stored pattern's offset bounds the table
            f.seek(t0)
            loc0 = 0
            while loc0 == 0:
                w = f.read(8)
                if len(w) < 8:
                    raise ValueError(f"{fp}: no stored pattern in the offset table")
                loc0 = int(np.frombuffer(w, dtype=np.uint64)[0])
            nslots = int(loc0 / 8 + 0.001)
            f.seek(t0)
            slots = np.frombuffer(f.read(8 * nslots), dtype=np.uint64)
            if self.version <= 4:
                # the count is the first slot index reached by the running minimum of the
                # table positions that the stored offsets point to
                pos = np.where(slots > 0, np.trunc((slots.astype(np.float64) - t0) / 8), np.inf)
                run = np.minimum.accumulate(np.concatenate(([int((loc0 - t0) / 8 + 0.001)], pos)))
                hit = np.flatnonzero(run[:len(pos)] <= np.arange(len(pos)))
                if hit.size == 0 or run[hit[0]] != hit[0]:
                    raise ValueError(f"{fp}: the offset table does not end where the first pattern starts")
                npat = int(hit[0])
            else:
                f.seek(loc0)
                ph = np.frombuffer(f.read(24), dtype=np.uint32)
                if ph[0] == 1:
                    raise NotImplementedError("compressed EBSP files are not supported")
                magic = int(ph[-3]) + (int(ph[-2]) << 32)
                hit = np.flatnonzero(slots == np.uint64(magic))
                if hit.size == 0:
                    raise ValueError(f"{fp}: no end-of-table marker for a version-5 file")
                npat = int(hit[0])
            self.nPatterns = np.int64(npat)
            f.seek(t0)
            self.filePos = np.frombuffer(f.read(8 * npat), dtype=np.uint64).copy()
            stored = self.filePos > 0
            if not stored.any():
                raise ValueError(f"{fp}: no stored pattern")
            first_rec = int(self.filePos[stored][0])
            hb = self._head_bytes()
            f.seek(first_rec)
            ph = np.frombuffer(f.read(hb), dtype=np.uint32)
            if ph[0] == 1:
                raise NotImplementedError("compressed EBSP files are not supported")
            self.patternW = np.uint32(ph[-2])
            self.patternH = np.uint32(ph[-3])
            nbytes = int(ph[-1])
            bits = nbytes / (int(self.patternW) * int(self.patternH)) * 8
            self.filedatatype = {8: np.uint8, 16: np.uint16, 32: np.uint32}.get(int(bits) if bits == int(bits) else 0)
            if self.filedatatype is None:
                raise ValueError(f"{fp}: {nbytes} bytes per {int(self.patternH)}x{int(self.patternW)} pattern")
            # positions: version 1 always stores them; from version 2 a flag byte says so
            flag = 0
            if self.version == 1:
                self.hasxypos = True
            else:
                f.seek(int(self.filePos[stored].min()) + hb + nbytes)
                b = f.read(1)
                self.hasxypos = len(b) == 1 and b[0] > 0
                flag = 1
            xall = np.zeros(npat, np.float64)
            yall = np.zeros(npat, np.float64)
            if self.hasxypos:
                # x at record + header + pixels (+ flag), y after x (+ flag)
                rec = self.filePos[stored].astype(np.int64)
                xo = rec + hb + nbytes + flag
                if rec.max() + hb + nbytes + 2 * flag + 16 > size:
                    raise ValueError(f"{fp}: a pattern record runs past the end of the file")
                mm = np.memmap(fp, dtype=np.uint8, mode='r')
                take = lambda o: np.ascontiguousarray(mm[o[:, None] + np.arange(8)]).view(np.float64).reshape(-1)
                xall[stored] = take(xo)
                yall[stored] = take(xo + 8 + flag)
                del mm
                self.xStep = xall[1] - xall[0] if npat > 1 else 0.0
                ncol = np.round((xall.max() - xall.min()) / self.xStep + 1) if self.xStep > 1e-6 else 1
                self.nCols = np.uint64(ncol)
                # the reference takes the row step from pattern nCols (one row down)
                self.yStep = np.abs(yall[0] - yall[int(self.nCols)]) if npat > int(self.nCols) else 0.0
                if self.yStep > 1e-6:
                    self.nRows = np.int64(np.round((yall.max() - yall.min()) / self.yStep + 1))
                else:
                    self.nRows = np.int64(npat / int(self.nCols) + 0.001)
            else:
                self.xStep = 1.0
                self.yStep = 1.0
                self.nCols = 1
                self.nRows = npat
            self._xpos, self._ypos = xall, yall
        if self.nCols is None or self.nCols == 0:
            self.nCols = np.uint64(1)
        if self.nRows is None:
            self.nRows = np.uint64(np.floor(self.nPatterns / self.nCols))
        return 0

    def _runs(self, patStart, n):
        pb = self.pattern_bytes
        pos = self.filePos[patStart:patStart + n].astype(np.int64)
        idx = np.flatnonzero(pos > 0)
        if idx.size == 0:
            return []
        off = pos[idx] + self._head_bytes()
        d = np.diff(off)
        stride = int(d[0]) if d.size else pb
        if stride < pb:
            stride = -1                       # overlapping records: one pattern per run
        brk = np.flatnonzero((np.diff(idx) != 1) | (d != stride)) + 1
        runs = []
        for seg in np.split(np.arange(idx.size), brk):
            runs.append((int(off[seg[0]]), stride if seg.size > 1 else pb, int(idx[seg[0]]), int(seg.size)))
        return runs

    def _xy(self, patStart, n):
        """The stored positions (float64; zeros where the file holds none), as the
        reference's pat_reader returns them (ebsd_pattern.py:910-957)."""
        xy = np.zeros((n, 2), np.float64)
        if self.hasxypos and self._xpos is not None:
            xy[:, 0] = self._xpos[patStart:patStart + n]
            xy[:, 1] = self._ypos[patStart:patStart + n]
        return xy

    def _footer_bytes(self):
        if self.version == 1:
            return 16
        return 18 if self.hasxypos else 1

    def _record_dtype(self):
        H, W = int(self.patternH), int(self.patternW)
        fields = [('head', np.uint32, (4,)), ('pat', np.dtype(self.filedatatype), (H, W))]
        if self.version == 1:
            fields += [('x', '<f8'), ('y', '<f8')]
        elif self.hasxypos:
            fields += [('fx', np.uint8), ('x', '<f8'), ('fy', np.uint8), ('y', '<f8')]
        else:
            fields += [('fx', np.uint8)]
        return np.dtype(fields)

    def _records(self, first, count, pats=None):
        """Records of patterns [first, first+count) in the reference writer's form: header
        (0, H, W, byte count), pixels, positions on the scan grid when the file has them."""
        H, W = int(self.patternH), int(self.patternW)
        rec = np.zeros(count, dtype=self._record_dtype())
        rec['head'] = (0, H, W, H * W * np.dtype(self.filedatatype).itemsize)
        if pats is not None:
            rec['pat'] = pats
        if self.version == 1 or self.hasxypos:
            i = np.arange(first, first + count)
            nc, nr = int(self.nCols), int(self.nRows)
            rec['x'] = ((i % nc) - nc * 0.5) * float(self.xStep)
            rec['y'] = ((i // nc) - nr * 0.5) * float(self.yStep)
            if self.version >= 2:
                rec['fx'] = 1
                rec['fy'] = 1
        return rec

    def write_header(self, writeBlank=False, bitdepth=8):
        """Version word, offset table and (``writeBlank``) zero patterns for the whole scan
        (ebsd_pattern.py:959-1053).  Versions 1-4 (version 2 unless set); the positions are
        written when ``hasxypos`` is set."""
        if bitdepth is None and self.filedatatype is None:
            raise ValueError('Error: extension not recognized, set "bitdepth" parameter')
        if bitdepth == 8:
            self.filedatatype = np.uint8
        elif bitdepth == 16:
            self.filedatatype = np.uint16
        if self.version is None:
            self.version = 2
        if not 1 <= int(self.version) <= 4:
            raise NotImplementedError(f"writing EBSP version {self.version}: versions 1 to 4 are written")
        if self.nPatterns is None:
            self.nPatterns = np.int64(int(self.nCols) * int(self.nRows))
        npat = int(self.nPatterns)
        t0 = self._table_start()
        if self.filePos is None:
            recbytes = 16 + self._footer_bytes() + self.pattern_bytes
            self.filePos = (np.arange(npat, dtype=np.uint64) * np.uint64(recbytes) + np.uint64(t0 + 8 * npat))
        fp = Path(self.filepath).expanduser()
        with open(fp, 'r+b' if fp.is_file() else 'w+b') as f:
            f.seek(0)
            f.write(np.array([-int(self.version)], np.int64).tobytes())
            if self.version >= 4:
                f.write(b'\0')
            f.write(np.asarray(self.filePos, dtype=np.uint64).tobytes())
            if writeBlank:
                rows = max(1, (64 << 20) // max(1, self._record_dtype().itemsize))
                for i0 in range(0, npat, rows):
                    cnt = min(rows, npat - i0)
                    f.write(self._records(i0, cnt).tobytes())

    def pat_writer(self, pat2write, patStart, nPatToWrite, typewrite=None):
        """Write patterns [patStart, patStart+n) into their records: header, pixels (as the
        file's type), and the scan-grid positions when the file has them
        (ebsd_pattern.py:1055-1088).  Records the table marks as not stored are skipped."""
        n = int(nPatToWrite)
        p0 = int(patStart)
        rec = self._records(p0, n, np.asarray(pat2write[:n]).astype(self.filedatatype))
        size = rec.dtype.itemsize
        # a version 2+ file without positions keeps its flag byte as write_header left it
        keep = size if (self.version == 1 or self.hasxypos) else 16 + self.pattern_bytes
        raw = rec.view(np.uint8).reshape(n, size)[:, :keep]
        pos = self.filePos[p0:p0 + n].astype(np.int64)
        with open(Path(self.filepath).expanduser(), 'r+b') as f:
            k = 0
            while k < n:
                if pos[k] <= 0:
                    k += 1
                    continue
                e = k + 1                     # records back to back: one write
                while keep == size and e < n and pos[e] == pos[e - 1] + size:
                    e += 1
                f.seek(int(pos[k]))
                f.write(raw[k:e].tobytes())
                k = e


class TFPatFile(EBSDPatternFile):
    """ThermoFisher .pat raw pattern file (ebsd_pattern.py:1090-1396)."""

    def __init__(self, path=None):
        EBSDPatternFile.__init__(self, path)
        self.filetype = 'TFPAT'
        self.vendor = 'THERMOFISHER'
        self.filedatatype = None
        self.bitdepth = 8
        self.filePos = None
        self.hasxypos = False
        self.hasbackground = True
        self.background = None
        self.isscan = True
        self.xypx = None
        self.xStep = 1.0
        self.yStep = 1.0

    @property
    def record_bytes(self):
        return 8 + int(self.patternW) * int(self.patternH) * 2

    def read_header(self, path=None, bitdepth=None):
        """The 25-byte header, the stored background and, from a ``.ebsd`` file beside
        it, the step (ebsd_pattern.py:1113-1196).  Returns 0, or -1 if the file cannot be
        opened."""
        if path is not None:
            self.set_filepath(path)
        fp = Path(self.filepath).expanduser()
        try:
            f = open(fp, 'rb')
        except OSError:
            print("File Not Found:", str(fp))
            return -1
        with f:
            head = f.read(25)
            size = f.seek(0, 2)
        version = struct.unpack_from('<I', head, 0)[0] if len(head) >= 4 else 0
        self.version = version
        if version < 1:
            raise ValueError(f"{fp}: TF .pat files of version 0 are not read (nor by the reference)")
        if version >= 1:
            self.filedatatype = np.uint16
            nc, nr, W, H = struct.unpack_from('<4I', head, 4)
            self.nCols, self.nRows = np.uint32(nc), np.uint32(nr)
            self.patternW, self.patternH = np.uint32(W), np.uint32(H)
            scanflag, = struct.unpack_from('<h', head, 20)
            self.bitdepth = np.uint8(head[22])
            backflag, = struct.unpack_from('<h', head, 23)
            self.hasbackground = backflag < 0
            self.filePos = 25 + (W * H * 2 if self.hasbackground else 0)
            if scanflag < 0:
                self.nPatterns = np.int64(nc) * np.int64(nr)
            else:
                self.nPatterns = np.int64(size - self.filePos) // np.int64(self.record_bytes)
        if self.isscan is False and self.nPatterns:
            self.xypx = self._locs(0, int(self.nPatterns))
        if self.hasbackground:
            self.background = self.read_backgroundpattern()
        self.read_stepsize()
        return 0

    def read_stepsize(self):
        """The step stored 45 bytes before the end of the ``.ebsd`` file beside this one,
        when it lies in (0, 2000) (ebsd_pattern.py:1198-1216)."""
        side = Path(str(self.filepath)).expanduser().with_suffix('.ebsd')
        try:
            with open(side, 'rb') as f:
                f.seek(-45, 2)
                step = np.frombuffer(f.read(8), dtype=np.float64)
        except OSError:
            return -1
        if step.size == 1 and 0 < step[0] < 2000:
            self.xStep = step[0]
            self.yStep = step[0]
        return 0

    def read_backgroundpattern(self):
        """The background pattern stored after the header (ebsd_pattern.py:1343-1360)."""
        if self.patternW is None or self.patternH is None:
            self.read_header()
        if not self.hasbackground:
            print('No background pattern stored.')
            return -1
        n = int(self.patternW) * int(self.patternH)
        with open(Path(self.filepath).expanduser(), 'rb') as f:
            f.seek(25)
            b = np.frombuffer(f.read(2 * n), dtype=np.uint16)
        return b.reshape(int(self.patternH), int(self.patternW)).copy()

    def get_backgroundpattern(self):
        return self.background

    def _runs(self, patStart, n):
        rb = self.record_bytes
        return [(int(self.filePos) + patStart * rb + 8, rb, 0, n)]

    def _locs(self, patStart, n):
        """The (column, row) words of records [patStart, patStart+n), uint32 (n, 2)."""
        mm = np.memmap(Path(self.filepath).expanduser(), dtype=np.uint8, mode='r')
        o = int(self.filePos) + (np.arange(patStart, patStart + n, dtype=np.int64) * self.record_bytes)
        loc = np.ascontiguousarray(mm[o[:, None] + np.arange(8)]).view(np.uint32).reshape(n, 2).copy()
        del mm
        return loc

    def _xy(self, patStart, n):
        """The records' pixel locations relative to the scan centre, times the step
        (ebsd_pattern.py:1218-1258), float32."""
        xy = self._locs(patStart, n).astype(np.float32)
        xy[:, 0] -= self.nCols * 0.5
        xy[:, 1] -= self.nRows * 0.5
        xy[:, 0] *= self.xStep
        xy[:, 1] *= self.yStep
        return xy

    def write_header(self, writeBlank=False, bitdepth=None):
        """The 25-byte header and the background pattern (``background`` when it has the
        pattern's size, else zeros); with ``writeBlank`` the file is extended to hold every
        record (ebsd_pattern.py:1260-1307)."""
        if bitdepth is not None:
            self.bitdepth = bitdepth
        if self.version is None:
            self.version = 1
        self.filedatatype = np.uint16
        W, H = int(self.patternW), int(self.patternH)
        head = struct.pack('<5IhBh', int(self.version), int(self.nCols), int(self.nRows), W, H,
                           0 if self.isscan is False else -1, int(self.bitdepth), -1 if self.hasbackground else 0)
        back = b''
        if self.hasbackground:
            bg = np.zeros((H, W), np.uint16)
            if isinstance(self.background, np.ndarray) and self.background.size == H * W:
                bg = self.background.reshape(H, W).astype(np.uint16)
            back = bg.tobytes()
        self.filePos = 25 + len(back)
        if self.nPatterns is None:
            self.nPatterns = np.int64(int(self.nCols) * int(self.nRows))
        fp = Path(self.filepath).expanduser()
        with open(fp, 'r+b' if fp.is_file() else 'w+b') as f:
            f.seek(0)
            f.write(head + back)
            if writeBlank:
                f.truncate(self.filePos + int(self.nPatterns) * self.record_bytes)

    def pat_writer(self, pat2write, patStart, nPatToWrite, typewrite=None, bitdepth=None):
        """Records for patterns [patStart, patStart+n): (column, row) on the scan grid and
        the pixels clipped to [0, 2^bitdepth - 1] (ebsd_pattern.py:1309-1341)."""
        if bitdepth is None:
            bitdepth = self.bitdepth
        n = int(nPatToWrite)
        H, W = int(self.patternH), int(self.patternW)
        top = int(2 ** int(bitdepth) - 1)
        rec = np.zeros(n, dtype=np.dtype([('loc', np.uint32, (2,)), ('pat', np.uint16, (H, W))]))
        i = np.arange(int(patStart), int(patStart) + n, dtype=np.int64)
        rec['loc'][:, 0] = i % int(self.nCols)
        rec['loc'][:, 1] = i // int(self.nCols)
        rec['pat'] = pat_flt2int(np.asarray(pat2write[:n]), typeout=np.uint16, method='clip', maxScale=top)
        with open(Path(self.filepath).expanduser(), 'r+b') as f:
            f.seek(int(self.filePos) + int(patStart) * self.record_bytes)
            f.write(rec.tobytes())
        return 0

    def copy_file(self, newpath, **kwargs):
        """Copy the file (``empty_data=True``: the header and background only, patterns
        zero) and the ``.ebsd`` file beside it (ebsd_pattern.py:1364-1396)."""
        if kwargs.get('empty_data'):
            dst = self._copy_header_only(newpath, int(self.filePos))
        else:
            dst = EBSDPatternFile.copy_file(self, newpath)
        side = Path(self.filepath).expanduser().resolve().with_suffix('.ebsd')
        if side.exists():
            shutil.copyfile(side, Path(dst).with_suffix('.ebsd'))
        return dst
