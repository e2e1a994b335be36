"""EDAX UP1/UP2 pattern files: the reference's reader interface, plus reads straight into
pinned host memory for the streaming pipeline.

Mirrors ``pyebsdindex.ebsd_pattern`` for the raw formats (``get_pattern_file_obj``
ebsd_pattern.py:33-116, ``EBSDPatterns`` :184-203, ``EBSDPatternFile.read_data`` :236-329,
``UPFile`` :435-708): the same header fields, the same ``read_data`` / ``pat_reader``
semantics and return values (patterns plus scan x/y locations), the same writer.

File layout (little-endian, as the reference reads it with ``np.fromfile``):

* version 1: uint32 version, uint32 W, uint32 H, uint32 data offset; patterns follow;
  the pattern count is (file size - 16) / (W * H * bytes per pixel) and the scan is one
  row of that many columns unless ``nCols`` was set.
* version >= 3: uint32 version, W, H, data offset; uint8 extra patterns; uint32 nCols,
  nRows; uint8 hex flag; float64 x step, y step (42 bytes); patterns follow at the offset.

Pixels are uint8 (``.up1``) or uint16 (``.up2``), row-major per pattern, patterns in scan
order.  Other formats of the reference (HDF5 families, .ebsp, .pat) are not built here:
the HDF5 ones need h5py, which this environment lacks.

``UPFile.read_into`` copies a run of patterns from the file into a caller's buffer (a
pinned host tensor of ``StreamingBandDetect``) with ``os.preadv`` on several threads:
the bytes go from the page cache into pinned memory once, with no intermediate array.
"""
import os
import struct
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np


def get_pattern_file_obj(path, file_type=str('')):
    """Pattern-file object for ``path``, its type from ``file_type`` or the extension, with
    its header read (ebsd_pattern.py:33-116).  Raises ValueError for an unknown extension
    and NotImplementedError for the reference's formats not built here."""
    filepath = np.atleast_1d(path)[0]
    ftype = file_type
    if ftype == str(''):
        ext = str.lower(Path(filepath).suffix)
        ftype = {'.up1': 'UP1', '.up2': 'UP2', '.pat': 'TFPAT', '.ebsp': 'EBSP', '.oh5': 'OH5',
                 '.h5': 'H5', '.h5oina': 'H5OINA', '.dm5': 'DM5'}.get(ext)
        if ftype is None:
            raise ValueError('Error: extension not recognized')
    if ftype.upper() in ('UP1', 'UP2'):
        obj = UPFile(path)
    else:
        raise NotImplementedError(f"pattern file type {ftype}: only EDAX UP1/UP2 files are read here")
    if obj.read_header() != 0:
        raise OSError(f"cannot read the header of {filepath}")
    return obj


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

    def pat_reader(self, patStart=0, nPatToRead=1):
        pass

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
            xyloc = np.zeros((ncol * nrow, 2), dtype=np.float32)
            for i in range(nrow):
                p0 = (row0 + i) * int(self.nCols) + col0
                pt, xy = self.read_data(convertToFloat=convertToFloat, patStartCount=[p0, ncol],
                                        returnArrayOnly=True)
                patterns[i * ncol:(i + 1) * ncol] = pt
                xyloc[i * ncol:(i + 1) * ncol] = xy
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
        """Scan locations of patterns [patStart, patStart+n) relative to the scan centre
        (ebsd_pattern.py:545-553)."""
        idx = np.arange(np.int64(patStart), np.int64(patStart + n), dtype=np.uint64)
        yx = np.unravel_index(idx, (np.int64(self.nRows), np.int64(self.nCols)))
        xy = np.array([yx[1], yx[0]]).T.copy().astype(np.float32)
        xy[:, 0] -= self.nCols * 0.5
        xy[:, 1] -= self.nRows * 0.5
        xy[:, 0] *= self.xStep
        xy[:, 1] *= self.yStep
        return xy


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

    def _pool(self, n):
        """A reader thread pool of at least ``n`` threads, kept between calls."""
        pool = getattr(self, "_readers", None)
        if pool is None or pool._max_workers < n:
            if pool is not None:
                pool.shutdown(wait=True)
            pool = self._readers = ThreadPoolExecutor(n)
        return pool

    @property
    def pattern_bytes(self):
        return int(self.patternW) * int(self.patternH) * np.dtype(self.filedatatype).itemsize

    def pat_reader(self, patStart=0, nPatToRead=1):
        """(patterns (n, H, W) of the file's type, xy (n, 2) float32) for patterns
        [patStart, patStart+n) (ebsd_pattern.py:515-555)."""
        out = np.empty((int(nPatToRead), int(self.patternH), int(self.patternW)), dtype=self.filedatatype)
        self.read_into(out, patStart, nPatToRead)
        return out, self._xyloc(patStart, int(nPatToRead))

    def read_into(self, dst, patStart, nPatToRead, threads=8):
        """Copy patterns [patStart, patStart+n) from the file into ``dst`` (a C-contiguous
        numpy array or CPU torch tensor, pinned or not, of at least n patterns of the file's
        size and type), with positioned reads on up to ``threads`` threads.  Raises
        ValueError for a short buffer, a wrong type or a range past the end of the file."""
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
        base = int(self.filePos) + int(patStart) * pb
        total = n * pb
        nthr = max(1, min(int(threads), total // (4 << 20) or 1))
        step = -(-total // nthr)
        fd = os.open(Path(self.filepath).expanduser(), os.O_RDONLY)
        try:
            def part(k):
                o, e = k * step, min(total, (k + 1) * step)
                while o < e:
                    got = os.preadv(fd, [view[o:e]], base + o)
                    if got <= 0:
                        raise OSError(f"short read at byte {base + o} of {self.filepath}")
                    o += got
            if nthr == 1:
                part(0)
            else:
                list(self._pool(nthr).map(part, range(nthr)))
        finally:
            os.close(fd)
        return dst

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
            if writeBlank:
                f.seek(int(self.filePos))
                blank = bytes(self.pattern_bytes)
                for _ in range(int(self.nRows) * int(self.nCols)):
                    f.write(blank)

    def pat_writer(self, pat2write, patStart, nPatToWrite, typewrite):
        """Write ``nPatToWrite`` patterns at pattern index ``patStart`` (ebsd_pattern.py:629-646)."""
        fp = Path(self.filepath).expanduser()
        nbytes = int(self.patternW) * int(self.patternH) * np.dtype(typewrite).itemsize
        with open(fp, 'r+b') as f:
            f.seek(nbytes * int(patStart) + int(self.filePos))
            np.ascontiguousarray(pat2write[0:nPatToWrite], dtype=typewrite).tofile(f)
