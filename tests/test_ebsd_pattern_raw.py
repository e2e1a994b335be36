"""Oxford .ebsp and ThermoFisher .pat files (pyebsdindex_amd/ebsd_pattern.py) against the
formats as the reference reads them (ebsd_pattern.py:709-1396).  No file of either format
ships with the reference, so files are assembled here field by field in the reference's
read order, with positions, steps and scan shapes worked out by hand; the writers are held
to the readers by round trips.  The version-5 EBSP layout follows the reference's reader
(the table slot that holds H + (W << 32)); no real version-5 file was available, so its
agreement with instrument files is unpinned."""
import numpy as np
import pytest

from pyebsdindex_amd import ebsd_pattern as ep


def _pats(n, H, W, dtype, seed=0, top=None):
    rng = np.random.default_rng(seed)
    hi = np.iinfo(dtype).max if top is None else top
    return rng.integers(0, hi, size=(n, H, W), endpoint=True).astype(dtype)


def _u32(*v):
    return np.array(v, np.uint32).tobytes()


def _f64(*v):
    return np.array(v, np.float64).tobytes()


def _ebsp(path, pats, xy=None, version=2, order=None, missing=(), head_words=None, marker=False):
    """An EBSP file: -version (int64), a mystery byte from version 4, the offset table,
    (version 5: an (H, W) word pair ending the table), then the records in ``order``: the
    per-pattern header, pixels and, with ``xy``, flag + x + flag + y (version 1: x + y)."""
    n, H, W = pats.shape
    nbytes = H * W * pats.dtype.itemsize
    t0 = 8 + (1 if version >= 4 else 0)
    hw = head_words if head_words is not None else ([0, H, W, nbytes] if version < 5 else [0, 0, 0, H, W, nbytes])
    rec_head = _u32(*hw)
    body = {}
    for i in range(n):
        r = rec_head + pats[i].tobytes()
        if version == 1:
            r += _f64(*(xy[i] if xy is not None else (0.0, 0.0)))
        elif xy is not None:
            r += b'\x01' + _f64(xy[i][0]) + b'\x01' + _f64(xy[i][1])
        else:
            r += b'\x00'
        body[i] = r
    order = list(range(n)) if order is None else list(order)
    pos = t0 + 8 * n + (8 if marker else 0)
    offs = np.zeros(n, np.uint64)
    data = b''
    for i in order:
        if i in missing:
            continue
        offs[i] = pos
        data += body[i]
        pos += len(body[i])
    out = np.array([-version], np.int64).tobytes() + (b'\x00' if version >= 4 else b'') + offs.tobytes()
    if marker:
        out += _u32(H, W)
    path.write_bytes(out + data)
    return offs


def _grid(ncols, nrows, xstep, ystep, x0=-3.0, y0=2.0):
    return [(x0 + (i % ncols) * xstep, y0 + (i // ncols) * ystep) for i in range(ncols * nrows)]


def test_ebsp_v2_header_by_hand(tmp_path):
    pats = _pats(6, 7, 9, np.uint16, seed=1)
    xy = _grid(3, 2, 0.5, 0.25)
    fn = tmp_path / "scan.ebsp"
    _ebsp(fn, pats, xy, version=2)
    f = ep.get_pattern_file_obj(str(fn))
    assert isinstance(f, ep.EBSPFile)
    assert (f.version, f.nPatterns, f.patternW, f.patternH) == (2, 6, 9, 7)
    assert f.filedatatype == np.uint16 and f.hasxypos
    # x step from the first two positions, columns from the x range: (1.0 / 0.5) + 1 = 3;
    # y step from pattern 0 against pattern nCols = 3, rows from the y range: 0.25 / 0.25 + 1
    assert (f.xStep, f.nCols) == (0.5, 3)
    assert (f.yStep, f.nRows) == (0.25, 2)
    got, pos = f.pat_reader(1, 4)
    assert np.array_equal(got, pats[1:5])
    assert pos.dtype == np.float64 and np.array_equal(pos, np.array(xy[1:5]))
    # the record table: 8 bytes of version, 6 offsets, then records of 16 + 126 + 18 bytes
    assert np.array_equal(f.filePos, 8 + 48 + np.arange(6) * (16 + 126 + 18))


def test_ebsp_unordered_records_and_missing_patterns(tmp_path):
    """Records in another order than the table, one pattern not stored (offset 0): the
    count is the table length, the missing pattern reads as zeros, the others as written."""
    pats = _pats(5, 6, 8, np.uint8, seed=2)
    xy = _grid(5, 1, 1.5, 0.0)
    fn = tmp_path / "holes.ebsp"
    offs = _ebsp(fn, pats, xy, version=3, order=[2, 0, 4, 1, 3], missing=(3,))
    f = ep.EBSPFile(fn)
    assert f.read_header() == 0
    assert f.nPatterns == 5 and offs[3] == 0
    got, pos = f.pat_reader(0, 5)
    want = pats.copy()
    want[3] = 0
    assert np.array_equal(got, want)
    assert np.array_equal(pos[[0, 1, 2, 4]], np.array(xy)[[0, 1, 2, 4]]) and np.all(pos[3] == 0)
    # one row: x step 1.5, 5 columns (pattern 3's position reads as 0 and is inside the range)
    assert (f.xStep, int(f.nCols), int(f.nRows)) == (1.5, 5, 1)
    # the reads are split into runs where the records are back to back
    runs = f._runs(0, 5)
    assert sum(r[3] for r in runs) == 4 and len(runs) >= 3
    buf = np.full((5, 6, 8), 7, np.uint8)
    f.read_into(buf, 0, 5)
    assert np.array_equal(buf, want)


@pytest.mark.parametrize("version,dtype,hasxy", [(1, np.uint8, True), (2, np.uint16, True), (2, np.uint8, False),
                                                 (3, np.uint16, True), (4, np.uint8, True), (4, np.uint16, False)])
def test_ebsp_writer_round_trip(tmp_path, version, dtype, hasxy):
    n_c, n_r = 4, 3
    pats = _pats(n_c * n_r, 10, 12, dtype, seed=version)
    fn = tmp_path / f"w{version}.ebsp"
    w = ep.EBSPFile(fn)
    w.version, w.hasxypos = version, hasxy
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = 12, 10, n_c, n_r, 0.5, 0.75
    w.nPatterns = n_c * n_r
    w.write_header(writeBlank=True, bitdepth=8 * np.dtype(dtype).itemsize)
    w.pat_writer(pats[3:9], 3, 6)
    r = ep.get_pattern_file_obj(fn)
    assert (r.version, int(r.nPatterns), int(r.patternW), int(r.patternH)) == (version, n_c * n_r, 12, 10)
    assert r.filedatatype == dtype
    got, pos = r.pat_reader(0, n_c * n_r)
    want = np.zeros_like(pats)
    want[3:9] = pats[3:9]
    assert np.array_equal(got, want)
    if hasxy or version == 1:
        assert r.hasxypos
        assert (r.xStep, int(r.nCols), r.yStep, int(r.nRows)) == (0.5, n_c, 0.75, n_r)
        i = np.arange(n_c * n_r)
        assert np.allclose(pos[:, 0], (i % n_c - n_c / 2) * 0.5) and np.allclose(pos[:, 1], (i // n_c - n_r / 2) * 0.75)
    else:
        assert not r.hasxypos
        assert (r.xStep, r.yStep, int(r.nCols), int(r.nRows)) == (1.0, 1.0, 1, n_c * n_r)
        assert np.all(pos == 0)
    # read_data of a slab (the scan shape as the reader found it)
    if hasxy:
        slab = r.read_data(patStartCount=[[1, 1], [2, 2]], returnArrayOnly=True)[0]
        assert np.array_equal(slab, want[[5, 6, 9, 10]])


def test_ebsp_version5_marker(tmp_path):
    pats = _pats(4, 5, 6, np.uint8, seed=5)
    xy = _grid(2, 2, 1.0, 2.0)
    fn = tmp_path / "v5.ebsp"
    _ebsp(fn, pats, xy, version=5, marker=True)
    f = ep.get_pattern_file_obj(fn)
    assert (f.version, int(f.nPatterns), int(f.patternW), int(f.patternH)) == (5, 4, 6, 5)
    assert (int(f.nCols), int(f.nRows), f.xStep, f.yStep) == (2, 2, 1.0, 2.0)
    got, pos = f.pat_reader(0, 4)
    assert np.array_equal(got, pats) and np.array_equal(pos, np.array(xy))


def test_ebsp_refusals(tmp_path):
    pats = _pats(3, 4, 4, np.uint8)
    fn = tmp_path / "c.ebsp"
    _ebsp(fn, pats, version=3, head_words=[1, 4, 4, 16])
    with pytest.raises(NotImplementedError, match="compressed"):
        ep.get_pattern_file_obj(fn)
    fn0 = tmp_path / "v0.ebsp"
    fn0.write_bytes(np.zeros(4, np.uint64).tobytes())
    with pytest.raises(ValueError):
        ep.get_pattern_file_obj(fn0)
    w = ep.EBSPFile(tmp_path / "v5w.ebsp")
    w.version, w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = 5, 4, 4, 2, 1, 1.0, 1.0
    with pytest.raises(NotImplementedError):
        w.write_header(writeBlank=True)


def test_ebsp_threaded_strided_reads(tmp_path):
    """A file large enough for several threads and vector reads of at most IOV_MAX / 2
    patterns each: every pattern lands in its place."""
    n = 1300
    pats = _pats(n, 60, 60, np.uint16, seed=9)
    fn = tmp_path / "big.ebsp"
    _ebsp(fn, pats, _grid(50, 26, 1.0, 1.0), version=2)
    f = ep.get_pattern_file_obj(fn)
    assert int(f.nPatterns) == n and (int(f.nCols), int(f.nRows)) == (50, 26)
    for start, cnt, thr in ((0, n, 4), (17, 1001, 3), (1299, 1, 8)):
        buf = np.empty((cnt, 60, 60), np.uint16)
        f.read_into(buf, start, cnt, threads=thr)
        assert np.array_equal(buf, pats[start:start + cnt]), (start, cnt, thr)
    with pytest.raises(ValueError):
        f.read_into(np.empty((2, 60, 60), np.uint16), n - 1, 2)


def _tfpat(path, pats, ncols, nrows, bitdepth=12, background=None, scan=True, tail=b'', locs=None):
    """A TF .pat file: the 25-byte header in the reference's read order, the background,
    then (column, row, pixels) records."""
    n, H, W = pats.shape
    head = (_u32(1, ncols, nrows, W, H) + np.array([-1 if scan else 0], np.int16).tobytes() +
            np.array([bitdepth], np.uint8).tobytes() +
            np.array([-1 if background is not None else 0], np.int16).tobytes())
    assert len(head) == 25
    body = b''
    for i in range(n):
        c, r = (i % ncols, i // ncols) if locs is None else locs[i]
        body += _u32(c, r) + pats[i].astype(np.uint16).tobytes()
    path.write_bytes(head + (background.astype(np.uint16).tobytes() if background is not None else b'') + body + tail)


def test_tfpat_header_by_hand(tmp_path):
    pats = _pats(12, 7, 5, np.uint16, seed=3, top=4095)
    bg = _pats(1, 7, 5, np.uint16, seed=4, top=4095)[0]
    fn = tmp_path / "map.pat"
    _tfpat(fn, pats, 4, 3, background=bg)
    # the step sits 45 bytes before the end of the .ebsd file beside it
    (tmp_path / "map.ebsd").write_bytes(bytes(100) + _f64(0.75) + bytes(37))
    f = ep.get_pattern_file_obj(str(fn))
    assert isinstance(f, ep.TFPatFile)
    assert (f.version, int(f.nCols), int(f.nRows), int(f.patternW), int(f.patternH)) == (1, 4, 3, 5, 7)
    assert (int(f.bitdepth), f.hasbackground, int(f.nPatterns), f.filePos) == (12, True, 12, 25 + 7 * 5 * 2)
    assert f.filedatatype == np.uint16
    assert np.array_equal(f.background, bg) and np.array_equal(f.get_backgroundpattern(), bg)
    assert (f.xStep, f.yStep) == (0.75, 0.75)
    got, xy = f.pat_reader(5, 4)
    assert np.array_equal(got, pats[5:9])
    # pattern 5 is column 1, row 1: ((1 - 2) * 0.75, (1 - 1.5) * 0.75); pattern 8: column 0, row 2
    assert xy.dtype == np.float32
    assert np.allclose(xy[0], [-0.75, -0.375]) and np.allclose(xy[3], [-1.5, 0.375])


def test_tfpat_count_from_size_and_step_range(tmp_path):
    """Scan flag >= 0: the count comes from the file size (a partial record at the end is
    not counted); no background: the records start at byte 25; a step outside (0, 2000) in
    the .ebsd file leaves the step at 1.  Stored locations, not the record index, give xy."""
    pats = _pats(5, 4, 6, np.uint16, seed=6)
    locs = [(9, 0), (8, 0), (7, 1), (6, 1), (5, 2)]
    fn = tmp_path / "pts.pat"
    _tfpat(fn, pats, 10, 3, scan=False, tail=bytes(11), locs=locs)
    (tmp_path / "pts.ebsd").write_bytes(bytes(10) + _f64(5000.0) + bytes(37))
    f = ep.get_pattern_file_obj(fn)
    assert (int(f.nPatterns), f.filePos, f.hasbackground, f.xStep) == (5, 25, False, 1.0)
    got, xy = f.pat_reader(0, 5)
    assert np.array_equal(got, pats)
    assert np.allclose(xy, np.array(locs, np.float32) - [5.0, 1.5])


def test_tfpat_writer_round_trip_and_copy(tmp_path):
    n_c, n_r = 3, 3
    pats = _pats(n_c * n_r, 8, 6, np.uint16, seed=7)             # values above 2^12 - 1 are clipped
    bg = np.arange(48, dtype=np.uint16).reshape(8, 6)
    fn = tmp_path / "w.pat"
    w = ep.TFPatFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows = 6, 8, n_c, n_r
    w.background = bg
    w.write_header(writeBlank=True, bitdepth=12)
    assert fn.stat().st_size == 25 + 96 + n_c * n_r * (8 + 96)
    w.pat_writer(pats[2:7], 2, 5)
    (tmp_path / "w.ebsd").write_bytes(bytes(8) + _f64(2.5) + bytes(37))
    r = ep.get_pattern_file_obj(fn)
    assert (int(r.nPatterns), int(r.bitdepth), r.xStep) == (n_c * n_r, 12, 2.5)
    assert np.array_equal(r.background, bg)
    got, xy = r.pat_reader(2, 5)
    assert np.array_equal(got, np.minimum(pats[2:7], 4095))
    i = np.arange(2, 7)
    assert np.allclose(xy, np.stack([(i % n_c - 1.5) * 2.5, (i // n_c - 1.5) * 2.5], 1))
    # copy_file(empty_data=True): header and background kept, patterns zero, .ebsd copied
    dst = r.copy_file(str(tmp_path / "empty.pat"), empty_data=True)
    c = ep.get_pattern_file_obj(dst)
    assert (int(c.nPatterns), c.xStep) == (n_c * n_r, 2.5) and np.array_equal(c.background, bg)
    assert np.all(c.pat_reader(0, n_c * n_r)[0] == 0)


def test_pat_flt2int_by_hand():
    x = np.array([[[-10.0, 0.0], [100.5, 300.0]]])
    assert np.array_equal(ep.pat_flt2int(x, np.uint8, 'clip'), [[[0, 0], [100, 255]]])
    assert ep.pat_flt2int(x).dtype == np.uint16                  # max > 258 picks uint16
    assert np.array_equal(ep.pat_flt2int(x, np.uint16, 'clip', maxScale=200), [[[0, 0], [100, 200]]])
    # fullscale: (x - min) * 0.98 * 255 / 310, rounded
    want = np.around((x - x.min()) * np.float32(0.98 * 255 / 310)).astype(np.uint8)
    assert np.array_equal(ep.pat_flt2int(x, np.uint8, 'fullscale'), want)
    # scale: lift the minimum to 0, then * 0.98 * 255 / max
    want = np.around((x + 10) * np.float32(0.98 * 255 / 310)).astype(np.uint8)
    assert np.array_equal(ep.pat_flt2int(x, np.uint8, 'scale'), want)
    assert ep.pat_flt2int(x, np.float32).dtype == np.float32


def test_write_data_runs_and_slabs(tmp_path):
    """The base class's write_data (ebsd_pattern.py:339-406) on a file that does not exist
    yet (header and blank patterns first), then a slab into it."""
    fn = tmp_path / "wd.up1"
    w = ep.UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = 5, 4, 4, 3, 1.0, 1.0
    w.nPatterns = 12
    run = np.full((3, 4, 5), 300.0)                              # clipped to 255 on the way in
    w.write_data(newpatterns=run, patStartCount=[2, -1])
    slab = _pats(4, 4, 5, np.uint8, seed=8)
    w.write_data(newpatterns=slab, patStartCount=[[2, 1], [2, 2]])
    r = ep.get_pattern_file_obj(fn)
    got = r.pat_reader(0, 12)[0]
    want = np.zeros((12, 4, 5), np.uint8)
    want[2:5] = 255
    want[[6, 7, 10, 11]] = slab
    assert np.array_equal(got, want)
