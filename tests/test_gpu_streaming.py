"""Pinned-host streaming: batches copied on a side stream and band-detected on another
must give exactly what find_bands gives on the same patterns resident on the device."""
import numpy as np
import pytest

from helpers import make_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,dtype,kw,sizes,batch", [
    ((60, 60), "uint8", {}, (16, 16, 7, 16, 3), 16),
    ((240, 240), "uint16", dict(nTheta=360, nRho=180, nBands=12), (4, 4, 2), 4),
])
def test_streaming_equals_resident(shape, dtype, kw, sizes, batch):
    import torch
    from pyebsdindex_amd.streaming import StreamingBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print(f"streaming {shape} {dtype} batches {sizes}", flush=True)
    bd, _ = make_pair(shape, **kw)
    fam = "HCP" if shape[0] == 240 else "FCC"
    pats = kikuchi_patterns(sum(sizes), shape, dtype, seed=51, family=fam)
    bd.backgroundsub = bd.backsub_fit(pats.numpy().astype(np.float64).mean(axis=0))
    want = {k: v.cpu().numpy() for k, v in bd.find_bands_device(pats.cuda()).items()}
    sb = StreamingBandDetect(bd, batch=batch, dtype=dtype, depth=2, chunk=batch)
    hb, off = [], 0
    for n in sizes:
        t = sb.pinned_batch(n)
        t.copy_(pats[off:off + n])
        hb.append(t)
        off += n
    got = sb.concat(sb.run(hb))
    torch.cuda.synchronize()
    for k in want:
        assert np.array_equal(got[k].numpy(), want[k]), k
    # a second pass through the same pipeline (slots reused after their events)
    got2 = sb.concat(sb.run(hb))
    for k in want:
        assert np.array_equal(got2[k].numpy(), want[k]), k


@pytest.mark.parametrize("shape,dtype,ext,start,count,batch,ring", [
    ((60, 60), "uint8", ".up1", 5, 50, 16, None),
    ((120, 120), "uint16", ".up2", 0, -1, 8, 2),
])
def test_file_streaming_equals_resident(tmp_path, shape, dtype, ext, start, count, batch, ring):
    """Patterns read from a UP1/UP2 file straight into the pinned ring and streamed give
    exactly what find_bands gives on the same patterns resident on the device (ragged last
    batch, an offset start, the smallest ring)."""
    import torch
    from pyebsdindex_amd.ebsd_pattern import UPFile, get_pattern_file_obj
    from pyebsdindex_amd.streaming import StreamingBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print(f"file streaming {shape} {dtype} start {start} count {count} batch {batch}", flush=True)
    n = 61
    pats = kikuchi_patterns(n, shape, dtype, seed=57).numpy()
    fn = tmp_path / f"scan{ext}"
    w = UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = shape[1], shape[0], n, 1, 1.0, 1.0
    w.write_header()
    w.pat_writer(pats, 0, n, pats.dtype)
    f = get_pattern_file_obj(str(fn))
    bd, _ = make_pair(shape)
    stop = n if count < 0 else start + count
    want = {k: v.cpu().numpy() for k, v in bd.find_bands_device(torch.from_numpy(pats[start:stop]).cuda()).items()}
    sb = StreamingBandDetect(bd, batch=batch, dtype=dtype, depth=2, chunk=batch)
    rbufs = None if ring is None else [sb.pinned_batch() for _ in range(ring)]
    got = sb.concat(sb.run_file(f, start, count, ring=rbufs, threads=3))
    for k in want:
        assert np.array_equal(got[k].numpy(), want[k]), k


def _write_raw(fn, pats, ncols, nrows):
    """The patterns as an .ebsp (version 3, positions) or .pat file of the package's writers."""
    from pyebsdindex_amd.ebsd_pattern import EBSPFile, TFPatFile
    n, H, W = pats.shape
    w = EBSPFile(fn) if str(fn).endswith(".ebsp") else TFPatFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = W, H, ncols, nrows, 1.0, 1.0
    w.nPatterns = n
    if isinstance(w, EBSPFile):
        w.version, w.hasxypos = 3, True
        w.write_header(writeBlank=True, bitdepth=8 * pats.dtype.itemsize)
    else:
        w.write_header(writeBlank=True, bitdepth=16)
    w.pat_writer(pats, 0, n)


@pytest.mark.parametrize("shape,dtype,ext,start,count,batch", [
    ((60, 60), "uint8", ".ebsp", 3, 50, 16),
    ((120, 120), "uint16", ".ebsp", 0, -1, 8),
    ((120, 120), "uint16", ".pat", 7, -1, 16),
])
def test_raw_file_streaming_equals_resident(tmp_path, shape, dtype, ext, start, count, batch):
    """Oxford .ebsp and ThermoFisher .pat files (records with headers and positions between
    the patterns) read into the pinned ring by vector reads: the bands equal find_bands on the
    same patterns resident on the device, bit for bit."""
    import torch
    from pyebsdindex_amd.ebsd_pattern import get_pattern_file_obj
    from pyebsdindex_amd.streaming import StreamingBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print(f"raw file streaming {ext} {shape} {dtype} start {start} count {count} batch {batch}", flush=True)
    n = 61
    pats = kikuchi_patterns(n, shape, dtype, seed=58).numpy()
    fn = tmp_path / f"scan{ext}"
    _write_raw(fn, pats, 61, 1)
    f = get_pattern_file_obj(str(fn))
    assert np.array_equal(f.pat_reader(0, n)[0], pats)
    bd, _ = make_pair(shape)
    stop = n if count < 0 else start + count
    want = {k: v.cpu().numpy() for k, v in bd.find_bands_device(torch.from_numpy(pats[start:stop]).cuda()).items()}
    sb = StreamingBandDetect(bd, batch=batch, dtype=dtype, depth=2, chunk=batch)
    got = sb.concat(sb.run_file(f, start, count, threads=3))
    for k in want:
        assert np.array_equal(got[k].numpy(), want[k]), k


def test_file_streaming_rejects_mismatches(tmp_path):
    """A file of another pattern type or size, a range past its end and a ring of one
    buffer are refused before anything is copied."""
    from pyebsdindex_amd.ebsd_pattern import UPFile, get_pattern_file_obj
    from pyebsdindex_amd.streaming import StreamingBandDetect
    bd, _ = make_pair((60, 60))
    sb = StreamingBandDetect(bd, batch=8, dtype="uint8", depth=2, chunk=8)
    fn = tmp_path / "s.up2"
    w = UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = 60, 60, 4, 1, 1.0, 1.0
    w.write_header(writeBlank=True)
    with pytest.raises(TypeError):
        sb.run_file(get_pattern_file_obj(str(fn)))
    fn8 = tmp_path / "s.up1"
    w8 = UPFile(fn8)
    w8.patternW, w8.patternH, w8.nCols, w8.nRows, w8.xStep, w8.yStep = 60, 60, 4, 1, 1.0, 1.0
    w8.write_header(writeBlank=True)
    f8 = get_pattern_file_obj(str(fn8))
    with pytest.raises(ValueError):
        sb.run_file(f8, 2, 5)
    with pytest.raises(ValueError):
        sb.run_file(f8, ring=[sb.pinned_batch()])
    fn9 = tmp_path / "t.up1"
    w9 = UPFile(fn9)
    w9.patternW, w9.patternH, w9.nCols, w9.nRows, w9.xStep, w9.yStep = 50, 60, 4, 1, 1.0, 1.0
    w9.write_header(writeBlank=True)
    with pytest.raises(ValueError):
        sb.run_file(get_pattern_file_obj(str(fn9)))


@pytest.mark.parametrize("shape,dtype,n,chunk,mode", [((120, 120), "uint16", 3000, -1, "fast"),
                                                      ((60, 60), "uint8", 1700, 600, None),
                                                      ((80, 80), "float32", 520, 256, "exact")])
def test_find_bands_api_pipeline_equals_device(shape, dtype, n, chunk, mode):
    """BandDetect.find_bands on host patterns (the call EBSDIndexer makes,
    _ebsd_index_single.py:703-707) runs them through the pinned-host pipeline; the
    structured array equals the device path's fields bit for bit (ragged last batch, a
    chunksize of the caller's, a background), and equals the direct path's (one upload)."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print(f"api pipeline {shape} {dtype} n={n} chunk={chunk}", flush=True)
    bd = BandDetect(patDim=shape, nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9,
                    background_mode=mode or "fast")
    pats = kikuchi_patterns(n, shape, dtype, seed=81).numpy()
    if mode:
        bd.collect_background(patsIn=pats[:200].astype(np.float64))
    got = bd.find_bands(pats, chunksize=chunk)
    assert getattr(bd, "_pipelines", None), "the host input did not take the pipelined path"
    dev = {k: v.cpu().numpy() for k, v in bd.find_bands_device(torch.from_numpy(pats).cuda()).items()}
    for k in ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid"):
        assert np.array_equal(got[k], dev[k]), k
    assert np.array_equal(bd.last_stats, dev["stats"])
    # the same call through the direct path (one upload of the whole array)
    bd.STREAM_MIN = n + 1
    direct = bd.find_bands(pats, chunksize=chunk)
    assert np.array_equal(direct, got)
    # a second pipelined call reuses the pipeline and its pinned buffers
    bd.STREAM_MIN = 512
    again = bd.find_bands(pats, chunksize=chunk)
    assert np.array_equal(again, got)


@pytest.mark.parametrize("n_phases", [1, 2, 10])
def test_pack_band_records_bytes(n_phases):
    """ebsd_pack_band_records writes the bands in the byte layout of BandDetect.dataType
    (band_detect.py:110-115): every byte equals numpy's own assembly of the structured array
    from the fields, for arbitrary float bit patterns (NaN payloads, -0, infinities) and
    int8 flags, with the fields find_bands does not compute left zero."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.band_detect import record_phases
    print(f"pack records n_phases={n_phases}", flush=True)
    bd = BandDetect(patDim=(60, 60), nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=7,
                    nPhases=n_phases)
    assert record_phases(bd.dataType) == n_phases
    plan = bd.device_plan()
    N = 37
    rng = np.random.default_rng(5 + n_phases)
    out = plan.alloc_out(N)
    host = {}
    for k, v in out.items():
        if k == "valid":
            a = rng.integers(-128, 128, size=tuple(v.shape), dtype=np.int8)
        else:
            a = rng.integers(0, 2**32, size=tuple(v.shape), dtype=np.uint32).view(np.float32)
        host[k] = a
        v.copy_(torch.from_numpy(a.copy()))
    rec = plan.pack_records(out, n_phases).cpu().numpy()
    want = np.zeros((N, bd.nBands), dtype=bd.dataType)
    for k in ("max", "normmax", "avemax", "maxloc", "aveloc", "width", "theta", "rho", "valid"):
        want[k] = host[k]
    assert rec.shape == (N * bd.nBands * bd.dataType.itemsize,)
    assert np.array_equal(rec, want.view(np.uint8).reshape(-1))


def test_find_bands_api_record_layouts():
    """The drop-in call with two phases (records of 81 bytes) and with a dataType of another
    layout (the per-field path): both equal the direct path's structured array."""
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.band_detect import record_phases
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print("api record layouts", flush=True)
    pats = kikuchi_patterns(700, (60, 60), "uint8", seed=83).numpy()
    for variant in ("two_phases", "other_layout"):
        bd = BandDetect(patDim=(60, 60), nTheta=180, nRho=90, tSigma=1, rSigma=2, rhoMaskFrac=0.1, nBands=9,
                        nPhases=2)
        if variant == "other_layout":
            bd.dataType = np.dtype([(n, bd.dataType.fields[n][0]) for n in bd.dataType.names], align=True)
            assert record_phases(bd.dataType) is None
        got = bd.find_bands(pats, chunksize=256)
        assert getattr(bd, "_pipelines", None), "the host input did not take the pipelined path"
        assert got.dtype == bd.dataType and got.shape == (700, 9)
        bd.PINNED_RESULT = False
        pageable = bd.find_bands(pats, chunksize=256)
        assert np.array_equal(pageable, got), variant
        bd.STREAM_MIN = 701
        direct = bd.find_bands(pats, chunksize=256)
        assert np.array_equal(direct, got), variant
        assert np.array_equal(got.view(np.uint8), direct.view(np.uint8)) or variant == "other_layout"
