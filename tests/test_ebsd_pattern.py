"""UP1/UP2 pattern files (pyebsdindex_amd/ebsd_pattern.py) against the format as the
reference reads it (ebsd_pattern.py:449-555): headers assembled here field by field in the
reference's read order, scan locations worked out by hand, slab reads against indexing."""
import numpy as np
import pytest

from pyebsdindex_amd import ebsd_pattern as ep


def _v3_file(path, pats, ncols, nrows, xstep=0.5, ystep=0.25, hexflag=0, extra=0, pos=42):
    """A version-3 file: the fields in the order the reference's np.fromfile calls read them."""
    H, W = pats.shape[1:]
    head = (np.array([3, W, H, pos], np.uint32).tobytes() + np.array([extra], np.uint8).tobytes() +
            np.array([ncols, nrows], np.uint32).tobytes() + np.array([hexflag], np.uint8).tobytes() +
            np.array([xstep, ystep], np.float64).tobytes())
    assert len(head) == 42
    with open(path, 'wb') as f:
        f.write(head + bytes(pos - 42) + np.ascontiguousarray(pats).tobytes())


def _pats(n, H, W, dtype, seed=0):
    rng = np.random.default_rng(seed)
    return rng.integers(0, np.iinfo(dtype).max, size=(n, H, W), endpoint=True).astype(dtype)


def test_v3_header_and_patterns(tmp_path):
    pats = _pats(12, 7, 9, np.uint16)
    fn = tmp_path / "scan.up2"
    _v3_file(fn, pats, ncols=4, nrows=3, hexflag=1, extra=2)
    f = ep.get_pattern_file_obj(str(fn))
    assert isinstance(f, ep.UPFile)
    assert (f.version, f.patternW, f.patternH, f.filePos) == (3, 9, 7, 42)
    assert (f.nCols, f.nRows, f.nPatterns, f.hexflag, f.extraPatterns) == (4, 3, 12, 1, 2)
    assert (f.xStep, f.yStep) == (0.5, 0.25)
    assert f.filedatatype == np.uint16
    got, xy = f.pat_reader(5, 4)
    assert np.array_equal(got, pats[5:9])
    # pattern 5 of a 4-column scan is column 1, row 1: x = (1 - 2) * 0.5, y = (1 - 1.5) * 0.25
    assert np.allclose(xy[0], [-0.5, -0.125])
    # pattern 8 is column 0, row 2
    assert np.allclose(xy[3], [-1.0, 0.125])


def test_v1_header_counts_from_file_size(tmp_path):
    pats = _pats(6, 5, 8, np.uint8, seed=1)
    fn = tmp_path / "old.up1"
    head = np.array([1, 8, 5, 16], np.uint32).tobytes()
    fn.write_bytes(head + pats.tobytes() + bytes(17))      # a partial pattern at the end is not counted
    f = ep.get_pattern_file_obj(fn)
    assert (f.version, f.nPatterns, f.nCols, f.nRows, f.xStep, f.yStep) == (1, 6, 6, 1, 0.0, 0.0)
    got, xy = f.pat_reader(0, 6)
    assert np.array_equal(got, pats)
    assert np.allclose(xy, 0.0)


def test_read_data_counts_and_slabs(tmp_path):
    pats = _pats(20, 4, 6, np.uint8, seed=2)
    fn = tmp_path / "s.up1"
    _v3_file(fn, pats, ncols=5, nrows=4)
    f = ep.UPFile(fn)
    assert f.read_header() == 0
    p, _ = f.read_data(patStartCount=[3, -1], returnArrayOnly=True)
    assert np.array_equal(p, pats[3:])
    p, _ = f.read_data(patStartCount=[7, 0], returnArrayOnly=True)         # count 0 reads one
    assert np.array_equal(p, pats[7:8])
    p, _ = f.read_data(patStartCount=[15, 100], returnArrayOnly=True)      # clipped to the scan
    assert np.array_equal(p, pats[15:])
    o = f.read_data(patStartCount=[[1, 1], [3, -1]], convertToFloat=True)  # columns 1-3, rows 1-3
    grid = pats.reshape(4, 5, 4, 6)
    assert o.patterns.dtype == np.float32
    assert np.array_equal(o.patterns, grid[1:4, 1:4].reshape(-1, 4, 6).astype(np.float32))
    assert list(o.nPatterns) == [3, 3] and list(o.patStart) == [1, 1]
    _, xy_run = f.read_data(patStartCount=[6, 3], returnArrayOnly=True)
    assert np.array_equal(o.xyLocations[:3], xy_run)


def test_read_into_threads_and_errors(tmp_path):
    """A read split over threads (file > 4 MiB per thread) is identical to one read; a
    wrong dtype, a short buffer or a range past the end raise."""
    pats = _pats(300, 120, 120, np.uint16, seed=3)                         # 8.6 MB
    fn = tmp_path / "big.up2"
    _v3_file(fn, pats, ncols=300, nrows=1)
    f = ep.get_pattern_file_obj(fn)
    a = np.empty((290, 120, 120), np.uint16)
    b = np.empty_like(a)
    f.read_into(a, 10, 290, threads=1)
    f.read_into(b, 10, 290, threads=8)
    assert np.array_equal(a, pats[10:]) and np.array_equal(b, a)
    import torch
    t = torch.empty((4, 120, 120), dtype=torch.uint16)
    f.read_into(t, 296, 4)
    assert np.array_equal(t.numpy(), pats[296:])
    with pytest.raises(ValueError):
        f.read_into(np.empty((4, 120, 120), np.uint8), 0, 4)
    with pytest.raises(ValueError):
        f.read_into(np.empty((3, 120, 120), np.uint16), 0, 4)
    with pytest.raises(ValueError):
        f.read_into(np.empty((4, 120, 120), np.uint16), 298, 4)


def test_writer_round_trip(tmp_path):
    pats = _pats(6, 10, 12, np.uint16, seed=4)
    fn = tmp_path / "w.up2"
    w = ep.UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = 12, 10, 3, 2, 1.5, 2.5
    w.write_header(writeBlank=True)
    assert fn.stat().st_size == 42 + 6 * 10 * 12 * 2
    w.pat_writer(pats[2:5], 2, 3, np.uint16)
    r = ep.get_pattern_file_obj(fn)
    assert (r.nCols, r.nRows, r.xStep, r.yStep) == (3, 2, 1.5, 2.5)
    got, _ = r.pat_reader(0, 6)
    want = np.zeros_like(pats)
    want[2:5] = pats[2:5]
    assert np.array_equal(got, want)


def test_unknown_and_unbuilt_types(tmp_path):
    with pytest.raises(ValueError):
        ep.get_pattern_file_obj(tmp_path / "x.tif")
    with pytest.raises(NotImplementedError):
        ep.get_pattern_file_obj(tmp_path / "x.h5")
