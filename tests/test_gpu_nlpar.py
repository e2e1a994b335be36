"""NLPAR on the device (libebsd_nlpar.so through its C ABI) against the oracle's C
restatement of the reference's Numba kernels (nlpar_cpu.py:753-935) on the same seeded
scans, plus the closed forms the oracle is pinned by, and the file-level NLPAR class.

Tolerances: the device sums each squared distance in float32 partials added in float64,
the reference sequentially in float32, so the distances differ by float32 rounding; the
normalised distance subtracts two nearly equal terms (d^2 and n (s_a^2 + s_b^2)), which
magnifies that rounding to about 1e-4 absolute.  Weights are compared to 2e-5 absolute,
averaged patterns to 1e-4 relative (+1e-3 absolute), integer outputs to 1 count on at
most 0.1% of the pixels (truncation of values that differ in the last bits)."""
import numpy as np
import pytest

from oracle import nlpar_ref as nr

pytestmark = pytest.mark.gpu


def _grain_scan(nrows, ncols, shape, dtype, seed, grains=4):
    """Scan of ``grains`` orientations in vertical stripes; every point is its grain's
    noise-free pattern with its own Poisson noise."""
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    clean = kikuchi_patterns(grains, shape, "float32", seed=seed,
                             background=2000.0 if dtype == "uint16" else 120.0,
                             amplitude=(300.0, 900.0) if dtype == "uint16" else (20.0, 60.0)).numpy()
    rng = np.random.default_rng(seed)
    gid = (np.arange(ncols)[None, :] * grains // ncols).repeat(nrows, 0).reshape(-1)
    pats = rng.poisson(np.maximum(clean[gid], 0)).astype(np.float64)
    if dtype == "float32":
        return pats.astype(np.float32)
    top = 255 if dtype == "uint8" else 65535
    return pats.clip(0, top).astype(dtype)


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("shape,dtype,nrows,ncols", [((60, 60), "uint8", 9, 11), ((40, 48), "uint16", 8, 7),
                                                     ((30, 30), "float32", 7, 9)])
def test_sigma_matches_oracle(shape, dtype, nrows, ncols):
    from pyebsdindex_amd.nlpar import makeautomask, sigma_device
    print(f"nlpar sigma {shape} {dtype} {nrows}x{ncols}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=21)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    s, d2, n2 = sigma_device(_dev(pats), nrows, ncols, idx, nn=1, return_d2=True)
    rs, rd2, rn2 = nr.sigma(pats.reshape(nrows * ncols, -1).astype(np.float32), nrows, ncols, idx, nn=1)
    assert np.allclose(s.cpu().numpy(), rs, rtol=1e-5)
    assert np.array_equal(n2.cpu().numpy(), rn2)
    assert np.allclose(d2.cpu().numpy(), rd2, rtol=1e-4, atol=2e-4)


@pytest.mark.parametrize("shape,dtype,nrows,ncols,sr,dthresh", [
    ((60, 60), "uint8", 9, 11, 3, 0.0), ((40, 48), "uint16", 8, 7, 2, 0.2), ((30, 30), "float32", 7, 9, 3, 0.0),
    ((33, 35), "uint8", 7, 8, 3, 0.0)])   # odd pattern size: no 4-pixel vector loads
def test_nlpar_matches_oracle(shape, dtype, nrows, ncols, sr, dthresh):
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device
    print(f"nlpar {shape} {dtype} {nrows}x{ncols} sr={sr}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=22)
    flat = pats.reshape(nrows * ncols, -1).astype(np.float32)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    rs = nr.sigma(flat, nrows, ncols, idx)[0]
    want, wref = nr.nlpar(flat, nrows, ncols, idx, rs, lam=0.7, sr=sr, dthresh=dthresh, return_weights=True)
    got, w = nlpar_device(_dev(pats), nrows, ncols, idx, _dev(rs), lam=0.7, sr=sr, dthresh=dthresh,
                          return_weights=True)
    assert np.allclose(w.cpu().numpy(), wref, atol=2e-5)
    assert np.allclose(got.cpu().numpy().reshape(want.shape), want, rtol=1e-4, atol=1e-3)
    # integer output as the reference writes it: clip, then truncate
    if dtype != "float32":
        import torch
        odt = {"uint8": torch.uint8, "uint16": torch.uint16}[dtype]
        gi = nlpar_device(_dev(pats), nrows, ncols, idx, _dev(rs), lam=0.7, sr=sr, dthresh=dthresh,
                          out_dtype=odt).cpu().numpy().reshape(want.shape).astype(np.int64)
        top = np.iinfo(dtype).max
        wi = want.clip(0, top).astype(dtype).astype(np.int64)
        diff = np.abs(gi - wi)
        assert diff.max() <= 1 and (diff > 0).mean() <= 1e-3


def test_closed_forms_on_device():
    """lambda -> infinity: the shifted-window mean; lambda -> 0 with small sigmas: the
    pattern itself."""
    from pyebsdindex_amd.nlpar import nlpar_device
    nrows, ncols, H, W, sr = 7, 9, 8, 8, 3
    rng = np.random.default_rng(23)
    data = (100 + 50 * rng.standard_normal((nrows * ncols, H, W))).astype(np.float32)
    idx = np.arange(H * W)
    sig = np.full((nrows, ncols), 0.05, np.float32)
    same = nlpar_device(_dev(data), nrows, ncols, idx, _dev(sig), lam=1e-3, sr=sr).cpu().numpy()
    assert np.array_equal(same, data)
    mean = nlpar_device(_dev(data), nrows, ncols, idx, _dev(sig), lam=1e6, sr=sr).cpu().numpy()
    cube = data.reshape(nrows, ncols, -1).astype(np.float64)
    ws = 2 * sr + 1
    for j in range(nrows):
        for i in range(ncols):
            y0, x0 = int(np.clip(j - sr, 0, nrows - ws)), int(np.clip(i - sr, 0, ncols - ws))
            want = cube[y0:y0 + ws, x0:x0 + ws].reshape(-1, H * W).mean(axis=0)
            assert np.allclose(mean[j * ncols + i].ravel(), want, rtol=2e-6)


def test_calclim_region_is_bit_identical():
    """A sub-region (the reference's calclim) gives bit for bit the points of the full run."""
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, sigma_device
    nrows, ncols, shape = 10, 12, (40, 40)
    pats = _grain_scan(nrows, ncols, shape, "uint8", seed=24)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    x = _dev(pats)
    s = sigma_device(x, nrows, ncols, idx)
    full = nlpar_device(x, nrows, ncols, idx, s).cpu().numpy().reshape(nrows, ncols, -1)
    part = nlpar_device(x, nrows, ncols, idx, s, calclim=(3, 2, 5, 6)).cpu().numpy().reshape(6, 5, -1)
    assert np.array_equal(part, full[2:8, 3:8])


@pytest.mark.parametrize("shape,dtype,nrows,ncols,sr,calclim,odt", [
    ((40, 40), "uint8", 11, 13, 3, None, "uint8"),
    ((48, 32), "uint16", 9, 10, 2, (1, 2, 7, 5), "float32"),     # odd calculation region
    ((32, 32), "float32", 8, 9, 1, (0, 1, 9, 7), "uint16"),
    ((40, 40), "uint8", 7, 7, 3, (2, 3, 3, 1), "float32"),       # one row of three points
])
def test_tiled_kernel_is_bit_identical(shape, dtype, nrows, ncols, sr, calclim, odt):
    """The 2 x 2 tiled NLPAR kernel (each window pattern loaded once per tile) gives the
    one-point-per-work-group kernel's averages and weights bit for bit: partial tiles at the
    region's edges, shifted windows at the scan's edges, every output type."""
    import torch
    from pyebsdindex_amd import _hip
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, sigma_device
    print(f"nlpar tiled {shape} {dtype} {nrows}x{ncols} sr={sr} calclim={calclim} out={odt}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=33)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    x = _dev(pats)
    s = sigma_device(x, nrows, ncols, idx)
    L = _hip.load_nlpar()
    res = []
    try:
        for tiled in (0, 1):
            L.ebsd_nlpar_set_tiled(tiled)
            o, w = nlpar_device(x, nrows, ncols, idx, s, sr=sr, calclim=calclim, out_dtype=getattr(torch, odt),
                                return_weights=True)
            res.append((o.cpu().numpy(), w.cpu().numpy()))
    finally:
        L.ebsd_nlpar_set_tiled(1)
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])


def test_nlpar_file_round_trip(tmp_path):
    """NLPAR.calcnlpar on a UP1 file equals the oracle on the whole scan (clipped to uint8),
    and a run in row bands equals the one-band run bit for bit; opt_lambda matches the same
    optimisation on the oracle's distances."""
    import scipy.optimize as opt
    from pyebsdindex_amd.ebsd_pattern import UPFile, get_pattern_file_obj
    from pyebsdindex_amd.nlpar import NLPAR, makeautomask
    nrows, ncols, shape = 12, 10, (50, 50)
    print("nlpar file round trip", flush=True)
    pats = _grain_scan(nrows, ncols, shape, "uint8", seed=25)
    fn = tmp_path / "scan.up1"
    w = UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = shape[1], shape[0], ncols, nrows, 1.0, 1.0
    w.write_header()
    w.pat_writer(pats, 0, nrows * ncols, np.uint8)

    nl = NLPAR(str(fn), lam=0.8, searchradius=3)
    out1 = nl.calcnlpar(fileout=str(tmp_path / "one.up1"))
    nl2 = NLPAR(str(fn), lam=0.8, searchradius=3, band_bytes=4 * ncols * 2500)   # bands of a few rows
    out2 = nl2.calcnlpar(fileout=str(tmp_path / "bands.up1"))
    a = get_pattern_file_obj(out1).pat_reader(0, nrows * ncols)[0]
    b = get_pattern_file_obj(out2).pat_reader(0, nrows * ncols)[0]
    assert np.array_equal(a, b)
    assert np.array_equal(nl.sigma, nl2.sigma)

    flat = pats.reshape(nrows * ncols, -1).astype(np.float32)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    rs, rd2, _ = nr.sigma(flat, nrows, ncols, idx)
    assert np.allclose(nl.sigma, rs, rtol=1e-5)
    want = nr.nlpar(flat, nrows, ncols, idx, rs, lam=0.8, sr=3).clip(0, 255).astype(np.uint8)
    diff = np.abs(a.reshape(want.shape).astype(int) - want.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() <= 1e-3

    lams = NLPAR(str(fn)).opt_lambda()

    def loptfunc(lam, d2, tw, dthresh):
        dw = np.exp(-np.maximum(d2, dthresh) / lam ** 2)
        return np.mean(np.abs(tw - 1.0 / (np.sum(dw, axis=2) + 1e-12)))
    ref = [opt.minimize(loptfunc, 1.0, args=(rd2, tw, np.float32(0.0)), method='Nelder-Mead',
                        bounds=[[0.001, 10.0]], options={'fatol': 0.0001})['x'][0] for tw in (0.5, 0.34, 0.25)]
    assert np.allclose(lams, ref, atol=2e-3)


@pytest.mark.parametrize("dtype,stem,backsub,rescale", [("uint8", True, False, False), ("uint16", False, True, False),
                                                       ("uint8", False, False, True), ("uint16", True, True, True)])
def test_nlpar_options_match_oracle(tmp_path, dtype, stem, backsub, rescale):
    """calcnlpar with stem_scale, backsub and rescale (nlpar_cpu.py:470-526) on the device
    against the oracle's restatement of the same data path around the C kernels: within one
    count on at most 0.1% of the pixels; with rescale, which multiplies the averages by
    top / range, within 1 + 2e-4 top (the averages' 1e-4 relative agreement, scaled)."""
    from pyebsdindex_amd.ebsd_pattern import UPFile, get_pattern_file_obj
    from pyebsdindex_amd.nlpar import NLPAR
    nrows, ncols, shape, sr, lam = 9, 10, (40, 40), 2, 0.8
    print(f"nlpar options {dtype} stem={stem} backsub={backsub} rescale={rescale}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=29)
    n = nrows * ncols
    fn = tmp_path / ("scan.up1" if dtype == "uint8" else "scan.up2")
    w = UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = shape[1], shape[0], ncols, nrows, 1.0, 1.0
    w.write_header()
    w.pat_writer(pats, 0, n, pats.dtype)
    nl = NLPAR(str(fn), lam=lam, searchradius=sr, stem_scale=stem)
    out = nl.calcnlpar(fileout=str(tmp_path / ("o.up1" if dtype == "uint8" else "o.up2")), backsub=backsub,
                       rescale=rescale)
    got = get_pattern_file_obj(out).pat_reader(0, n)[0].astype(np.int64)
    top = np.iinfo(dtype).max
    mask = nr.makeautomask(*shape)
    want = nr.calcnlpar_options(pats, nrows, ncols, np.flatnonzero(mask.ravel()), lam=lam, sr=sr, stem_scale=stem,
                                backsub=backsub, rescale_top=top if rescale else None, out_dtype=np.dtype(dtype),
                                mask=mask).reshape(got.shape).astype(np.int64)
    diff = np.abs(got - want)
    print(f"  max diff {diff.max()}, pixels differing {(diff > 0).mean():.5f}", flush=True)
    if rescale:
        assert diff.max() <= 1 + 2e-4 * top
    else:
        assert diff.max() <= 1 and (diff > 0).mean() <= 1e-3
    assert np.array_equal(nl.mask, mask)


@pytest.mark.parametrize("ext,dtype", [(".ebsp", "uint8"), (".pat", "uint16")])
def test_nlpar_raw_files_equal_up(tmp_path, ext, dtype):
    """NLPAR.calcnlpar on an .ebsp / .pat file writes the same patterns as on the same scan
    in an UP file (the output is a copy of the input with its patterns replaced,
    nlpar_cpu.py:590,603-613), and the copy keeps the input's scan layout."""
    from pyebsdindex_amd.ebsd_pattern import EBSPFile, TFPatFile, UPFile, get_pattern_file_obj
    from pyebsdindex_amd.nlpar import NLPAR
    nrows, ncols, shape = 9, 8, (40, 40)
    print(f"nlpar raw file {ext} {dtype}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=27)
    n = nrows * ncols
    up = tmp_path / ("scan.up1" if dtype == "uint8" else "scan.up2")
    w = UPFile(up)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = shape[1], shape[0], ncols, nrows, 1.0, 1.0
    w.write_header()
    w.pat_writer(pats, 0, n, pats.dtype)
    raw = tmp_path / f"scan{ext}"
    r = EBSPFile(raw) if ext == ".ebsp" else TFPatFile(raw)
    r.patternW, r.patternH, r.nCols, r.nRows, r.xStep, r.yStep = shape[1], shape[0], ncols, nrows, 1.0, 1.0
    r.nPatterns = n
    if ext == ".ebsp":
        r.version, r.hasxypos = 2, True
        r.write_header(writeBlank=True, bitdepth=8)
    else:
        r.write_header(writeBlank=True, bitdepth=16)
    r.pat_writer(pats, 0, n)
    a = get_pattern_file_obj(NLPAR(str(up), lam=0.9, searchradius=2).calcnlpar(
        fileout=str(tmp_path / ("o.up1" if dtype == "uint8" else "o.up2")))).pat_reader(0, n)[0]
    out = NLPAR(str(raw), lam=0.9, searchradius=2).calcnlpar(fileout=str(tmp_path / f"o{ext}"))
    fo = get_pattern_file_obj(out)
    b = fo.pat_reader(0, n)[0]
    assert type(fo) is type(r) and (int(fo.nCols), int(fo.nRows)) == (ncols, nrows)
    assert np.array_equal(a, b)
    assert not np.array_equal(b, pats)


@pytest.mark.parametrize("dtype", ["uint8", "uint16"])
def test_saturated_patterns_and_both_paths(dtype):
    """Integer scans with saturated pixels in some patterns (1 to 60): the packed-word path
    (with the saturated pixels' terms taken out, or the per-pixel path for patterns with
    more than 32) and the per-pixel path both match the oracle."""
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, sigma_device
    nrows, ncols, shape = 9, 10, (40, 40)
    print(f"nlpar saturated {dtype}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=26)
    top = np.iinfo(dtype).max
    rng = np.random.default_rng(26)
    # 1-32 saturated pixels: exact dot products with the saturated terms taken out; more
    # than 32: per-pixel path
    for k, p in enumerate(rng.choice(nrows * ncols, 16, replace=False)):
        nsat = (1, 7, 32, 33, 60)[k % 5]
        pats[p].reshape(-1)[rng.choice(shape[0] * shape[1], nsat, replace=False)] = top
    flat = pats.reshape(nrows * ncols, -1).astype(np.float32)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    rs = nr.sigma(flat, nrows, ncols, idx)[0]
    want = nr.nlpar(flat, nrows, ncols, idx, rs, lam=0.7, sr=3)
    x = _dev(pats)
    for fast in (True, False):
        s, d2, n2 = sigma_device(x, nrows, ncols, idx, return_d2=True, fast=fast)
        assert np.allclose(s.cpu().numpy(), rs, rtol=1e-5), fast
        got = nlpar_device(x, nrows, ncols, idx, _dev(rs), lam=0.7, sr=3, fast=fast).cpu().numpy()
        assert np.allclose(got.reshape(want.shape), want, rtol=1e-4, atol=1e-3), fast


@pytest.mark.parametrize("dtype,n,P", [("uint8", 37, 61), ("uint16", 5, 3600), ("float32", 9, 13)])
def test_pattern_max(dtype, n, P):
    """The device maximum equals numpy's, with ragged tails after the 16-value loads and
    negative float values."""
    import torch
    from pyebsdindex_amd.nlpar import pattern_max
    rng = np.random.default_rng(27)
    if dtype == "float32":
        x = (rng.standard_normal((n, P)) * 100 - 300).astype(np.float32)
    else:
        x = rng.integers(0, np.iinfo(dtype).max // 3, size=(n, P)).astype(dtype)
    x.reshape(-1)[-1] = x.max() + (1 if dtype != "float32" else 0.5)   # the maximum in the tail
    assert pattern_max(_dev(x)) == float(x.max())
    # an unaligned start (offset by one element) takes the element-by-element path
    flat = _dev(x.reshape(-1))
    assert pattern_max(flat[1:].reshape(1, -1)) == float(x.reshape(-1)[1:].max())


@pytest.mark.parametrize("shape,dtype,nrows,ncols,nn,sr", [
    ((60, 60), "uint8", 11, 12, 2, 1), ((48, 48), "uint8", 12, 11, 1, 5)])
def test_other_radii(shape, dtype, nrows, ncols, nn, sr):
    """Neighbourhoods of +-2 for sigma and search radii 1 and 5 (windows of 9 and 121)
    against the oracle's restatement."""
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, sigma_device
    print(f"nlpar radii {shape} {dtype} {nrows}x{ncols} nn={nn} sr={sr}", flush=True)
    pats = _grain_scan(nrows, ncols, shape, dtype, seed=29)
    flat = pats.reshape(nrows * ncols, -1).astype(np.float32)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    x = _dev(pats)
    s, d2, n2 = sigma_device(x, nrows, ncols, idx, nn=nn, return_d2=True)
    rs, rd2, rn2 = nr.sigma(flat, nrows, ncols, idx, nn=nn)
    assert np.allclose(s.cpu().numpy(), rs, rtol=1e-5)
    assert np.array_equal(n2.cpu().numpy(), rn2)
    assert np.allclose(d2.cpu().numpy(), rd2, rtol=1e-4, atol=2e-4)
    want, wref = nr.nlpar(flat, nrows, ncols, idx, rs, lam=0.7, sr=sr, return_weights=True)
    got, w = nlpar_device(x, nrows, ncols, idx, _dev(rs), lam=0.7, sr=sr, return_weights=True)
    assert np.allclose(w.cpu().numpy(), wref, atol=2e-5)
    assert np.allclose(got.cpu().numpy().reshape(want.shape), want, rtol=1e-4, atol=1e-3)


def test_large_patterns_against_float64():
    """240x240 uint16 (43,433 mask pixels): the reference's sequential float32 sums drift
    from the exact distances (the restatement's sigma is 3.7e-5 off the float64 one here),
    so the device, whose integer sums are exact, is held to the float64 oracle: sigma within
    1e-6, the averaged patterns (with the float64 sigma) within 5e-5."""
    from pyebsdindex_amd.nlpar import makeautomask, nlpar_device, sigma_device
    nrows, ncols, shape = 7, 8, (240, 240)
    print("nlpar 240x240 uint16 vs float64", flush=True)
    pats = _grain_scan(nrows, ncols, shape, "uint16", seed=29)
    flat = pats.reshape(nrows * ncols, -1).astype(np.float64)
    idx = np.nonzero(makeautomask(*shape).ravel())[0]
    x = _dev(pats)
    s64 = nr.sigma64(flat, nrows, ncols, idx)
    s = sigma_device(x, nrows, ncols, idx).cpu().numpy()
    assert np.allclose(s, s64, rtol=1e-6)
    sig = s64.astype(np.float32)
    want = nr.nlpar64(flat, nrows, ncols, idx, sig, lam=0.7, sr=3)
    got = nlpar_device(x, nrows, ncols, idx, _dev(sig), lam=0.7, sr=3).cpu().numpy()
    assert np.allclose(got.reshape(want.shape), want, rtol=5e-5, atol=1e-3)
