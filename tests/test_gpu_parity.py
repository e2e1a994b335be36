"""Device parity: HIP kernels (through the C ABI) against the oracle on the same seeded inputs."""
import numpy as np
import pytest

from helpers import compare_bands, make_pair, oracle64
from oracle import fp64

pytestmark = pytest.mark.gpu


def _pats(n, shape, dtype, seed, **kw):
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    return kikuchi_patterns(n, shape, dtype, seed=seed, **kw).numpy()


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("shape,dtype,n", [((60, 60), "uint8", 37), ((120, 120), "uint16", 19),
                                           ((80, 80), "uint8", 5), ((100, 120), "uint16", 3),
                                           ((60, 60), "float32", 9)])
def test_radon_matches_oracle_bitwise(shape, dtype, n):
    """u8/u16: integer line sums are exact, so float32(sum/count) is bit-identical to the
    reference's float64 accumulation (radon_fast.py:288-310).  float32 input: the same
    float64 sum in the same (sorted-plan) order -> also bit-identical."""
    print(f"radon {shape} {dtype} n={n}", flush=True)
    bd, ref = make_pair(shape)
    pats = _pats(n, shape, dtype, seed=3)
    rad = bd.device_plan().radon(_dev(pats)).cpu().numpy()
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    assert rad.shape == want.shape
    assert np.array_equal(rad, want)


@pytest.mark.parametrize("dtype", ["uint8", "uint16"])
def test_radon_every_pack_size(dtype):
    """Every patterns-per-word variant of the integer gather (u8: 2, 4, 8 per word; u16: 1, 2,
    4, 8), each with its own accumulation layout, is bit-identical to the reference
    restatement, with and without row segments; 13 patterns leave a ragged last pack."""
    print(f"every pack size {dtype}", flush=True)
    shape = (60, 60)
    bd, ref = make_pair(shape)
    pats = _pats(13, shape, dtype, seed=53)
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    plan = bd.device_plan()
    x = _dev(pats)
    for gp in (1, 2, 4, 8):
        plan.set_tuning(gather_pack=gp)
        for S in (1, 2):
            plan.set_row_segments(S)
            assert np.array_equal(plan.radon(x).cpu().numpy(), want), (gp, S)
    plan.set_tuning(gather_pack=4)
    plan.set_row_segments(0)


@pytest.mark.parametrize("vmax", [65535, 8192, 8191, 4095, 1023, 255, 1, 0])
def test_radon_u16_value_ranges(vmax):
    """uint16 sums per pack: the wide form (a pack holding a value above 8,191) and the narrow
    form with every flush interval (8,191: every group, 4,095: every two, ..., 0: only at
    line ends), for each patterns-per-word variant and row segmentation, with one pack of
    large values among small ones -- bit-identical to the reference restatement."""
    print(f"u16 value range {vmax}", flush=True)
    shape = (80, 80)
    bd, ref = make_pair(shape)
    x = _pats(11, shape, "uint16", seed=57).astype(np.float64)
    x = np.rint(x / max(x.max(), 1.0) * vmax).astype(np.uint16)
    x[3, 40, 41] = vmax          # the largest value exactly, in one pattern
    x[9] = np.minimum(x[9].astype(np.int64) * 8, 65535).astype(np.uint16)  # a pack of larger values
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(x, ref.padding)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    plan = bd.device_plan()
    xd = _dev(x)
    for gp in (2, 4, 8):
        plan.set_tuning(gather_pack=gp)
        for S in (1, 2):
            plan.set_row_segments(S)
            assert np.array_equal(plan.radon(xd).cpu().numpy(), want), (gp, S)
    plan.set_tuning(gather_pack=4)
    plan.set_row_segments(0)


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_radon_background(mode):
    print(f"radon background {mode}", flush=True)
    shape = (120, 120)
    bd, ref = make_pair(shape, background_mode=mode)
    pats = _pats(8, shape, "uint16", seed=4)
    bg = bd.backsub_fit(pats.mean(axis=0))
    bd.backgroundsub = bg
    plan = bd.device_plan()
    x, flags = bd._prepare(plan, _dev(pats))
    rad = plan.radon(x, flags=flags).cpu().numpy()
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding, background=bg)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    if mode == "exact":
        assert np.array_equal(rad, want)
    else:
        r64 = fp64.radon64(pats, ref.radonPlan.indexPlan, bg)
        scale = np.abs(r64).max(axis=(1, 2), keepdims=True)
        assert np.abs(rad - r64).max() <= 1e-6 * scale.max()
        assert (np.abs(rad - want) / scale).max() <= 1e-5


def test_conv_matches_fp64():
    """Separable float32 convolution vs the FP64 direct sum with the exact 9x5 kernel:
    within 1e-5 of max|C| (the reference's own float32 FFT is ~5e-7)."""
    shape = (120, 120)
    bd, ref = make_pair(shape)
    pats = _pats(6, shape, "uint16", seed=5)
    plan = bd.device_plan()
    rad = plan.radon(_dev(pats))
    conv = plan.conv(rad).cpu().numpy()
    p0, p1 = ref.padding
    C64 = fp64.conv64(fp64.pad64(rad.cpu().numpy().astype(np.float64), p0, p1), ref.kernel[0])
    err = np.abs(conv - C64).max(axis=(1, 2)) / np.abs(C64).max(axis=(1, 2))
    assert err.max() <= 1e-5, err


@pytest.mark.parametrize("shape,dtype,n,seed", [((60, 60), "uint8", 40, 11), ((120, 120), "uint16", 24, 12),
                                                ((160, 160), "uint16", 6, 13), ((100, 120), "uint8", 5, 14)])
def test_find_bands_vs_fp64(shape, dtype, n, seed):
    print(f"find_bands {shape} {dtype} n={n}", flush=True)
    bd, ref = make_pair(shape)
    pats = _pats(n, shape, dtype, seed=seed)
    got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(_dev(pats)).items()}
    res, st, C64, _ = oracle64(ref, pats)
    total, ties = compare_bands(got, res, C64, ref.padding, ref.peakPad)
    assert ties <= max(1, total // 100)
    np.testing.assert_allclose(got['stats'][:, 0], st[0], rtol=0, atol=1e-5 * np.abs(st[0]).max())
    np.testing.assert_allclose(got['stats'][:, 1], st[1], rtol=0, atol=1e-5 * np.abs(st[2]).max())
    np.testing.assert_allclose(got['stats'][:, 2], st[2], rtol=1e-5)
    same = (got['valid'] == 1) & (res['valid'] == 1) & np.all(got['maxloc'] == res['maxloc'], -1)
    np.testing.assert_allclose(got['max'][same], res['max'][same], rtol=2e-5, atol=1e-5 * np.abs(res['max']).max())
    np.testing.assert_allclose(got['theta'][same], res['theta'][same], atol=1e-3 * np.pi / 180)
    np.testing.assert_allclose(got['rho'][same], res['rho'][same], atol=1e-3 * 2 * bd.rhoMax / bd.nRho)
    np.testing.assert_allclose(got['width'][same], res['width'][same], rtol=1e-3, atol=1e-4)


@pytest.mark.parametrize("case,mode", [("120-u16", "fast"), ("120-u16", "exact"),
                                       ("240-360x180", "fast"), ("240-360x180", "exact")])
def test_find_bands_background_vs_fp64(case, mode):
    """Whole find_bands with background subtraction -- the benchmarked configurations: 120x120
    uint16 FCC with a 180x90 Radon, and 240x240 uint16 HCP (c/a 1.587) with a 360x180 Radon and
    12 bands -- against the FP64 oracle given the same background (x - b per pixel in float64,
    radon_fast.py:255-259): integer peak bins identical except FP64 near-ties, sub-pixel
    positions within 1e-3 bin, stats within 1e-5.  The exact mode is also held to the
    reference restatement, which rounds every x - b to float32 as the reference does: the
    Radon bit for bit, and its bands."""
    print(f"find_bands background {case} {mode}", flush=True)
    if case == "120-u16":
        shape, kw, n, fam = (120, 120), {}, 24, "FCC"
    else:
        shape, kw, n, fam = (240, 240), dict(nTheta=360, nRho=180, nBands=12), 5, "HCP"
    bd, ref = make_pair(shape, background_mode=mode, **kw)
    pats = _pats(n, shape, "uint16", seed=61, family=fam)
    bd.collect_background(patsIn=pats.astype(np.float64))
    bg = bd.backgroundsub
    assert bg is not None and bg.shape == shape
    got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(_dev(pats)).items()}
    res, st, C64, R64 = oracle64(ref, pats, background=bg)
    total, ties = compare_bands(got, res, C64, ref.padding, ref.peakPad)
    assert ties <= max(1, total // 100)
    assert (got['valid'] == 1).mean() >= 0.9
    np.testing.assert_allclose(got['stats'][:, 0], st[0], rtol=0, atol=1e-5 * np.abs(st[0]).max())
    np.testing.assert_allclose(got['stats'][:, 1], st[1], rtol=0, atol=1e-5 * np.abs(st[2]).max())
    np.testing.assert_allclose(got['stats'][:, 2], st[2], rtol=1e-5)
    same = (got['valid'] == 1) & (res['valid'] == 1) & np.all(got['maxloc'] == res['maxloc'], -1)
    np.testing.assert_allclose(got['max'][same], res['max'][same], rtol=2e-5, atol=1e-5 * np.abs(res['max']).max())
    np.testing.assert_allclose(got['theta'][same], res['theta'][same], atol=1e-3 * np.pi / 180)
    np.testing.assert_allclose(got['rho'][same], res['rho'][same], atol=1e-3 * 2 * bd.rhoMax / bd.nRho)
    # the device Radon itself: fast within 1e-6 of FP64, exact bit-identical to the restatement
    plan = bd.device_plan()
    x, flags = bd._prepare(plan, _dev(pats))
    rad = plan.radon(x, flags=flags).cpu().numpy()
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding, background=bg)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    if mode == "exact":
        assert np.array_equal(rad, want)
        ref.backgroundsub = bg
        rb = ref.find_bands(pats)
        agree = np.all(got['maxloc'] == rb['maxloc'], -1) & (got['valid'] == rb['valid'])
        assert agree.mean() >= 0.98
        np.testing.assert_allclose(got['aveloc'][agree], rb['aveloc'][agree], atol=1e-3)
    else:
        assert np.abs(rad - R64).max() <= 1e-6 * np.abs(R64).max()


def test_find_bands_api_matches_reference_restatement(al_sim_pattern):
    """BandDetect.find_bands (numpy in, structured array out) vs the reference-faithful CPU
    restatement on the reference's own test pattern, plus synthetic 8-bit patterns."""
    shape = al_sim_pattern.shape
    bd, ref = make_pair(shape)
    pats = np.concatenate([al_sim_pattern[None], _pats(7, shape, "uint8", seed=21)])
    got = bd.find_bands(pats)
    want = ref.find_bands(pats)
    assert got.dtype == want.dtype
    res, st, C64, _ = oracle64(ref, pats)
    total, ties = compare_bands({k: got[k] for k in ('valid', 'maxloc', 'aveloc')}, res, C64,
                                ref.padding, ref.peakPad)
    assert ties <= 1
    agree = np.all(got['maxloc'] == want['maxloc'], -1) & (got['valid'] == want['valid'])
    assert agree.mean() >= 0.98
    np.testing.assert_allclose(got['aveloc'][agree], want['aveloc'][agree], atol=1e-3)
    np.testing.assert_array_equal(got['band_match_index'], want['band_match_index'])


def test_edge_cases_constant_and_single():
    """Plateaus (constant patterns make every masked cell a local max), N=1, N=0."""
    import torch
    shape = (60, 60)
    bd, ref = make_pair(shape)
    const = np.full((2,) + shape, 77, dtype=np.uint8)
    got = bd.find_bands(const)
    want = ref.find_bands(const)
    assert np.array_equal(got['valid'], want['valid'])
    # every masked cell ties exactly (FP64 and the device's separable float32 alike), which
    # overflows the candidate list and exercises the chunked fallback; ties resolve to the
    # highest row-major positions in both
    res, st, C64, _ = oracle64(ref, const)
    assert np.array_equal(got['maxloc'], res['maxloc'].astype(np.float32))
    one = _pats(1, shape, "uint8", seed=8)
    g1 = bd.find_bands(one[0])
    assert g1.shape == (1, 9)
    res, st, C64, _ = oracle64(ref, one)
    compare_bands({k: g1[k] for k in ('valid', 'maxloc', 'aveloc')}, res, C64, ref.padding, ref.peakPad)
    out = bd.find_bands_device(torch.empty((0,) + shape, dtype=torch.uint8, device="cuda"))
    assert out['max'].shape == (0, 9)


@pytest.mark.parametrize("dtype", ["uint16", "uint8"])
def test_scan_scale_geometry_240(dtype):
    """240x240 patterns with a 360x180 Radon and 12 bands (the streaming workload's shape):
    2-byte LDS words in the gather and a 218x382 padded Radon in the peak kernel."""
    print(f"240x240 {dtype}", flush=True)
    shape = (240, 240)
    bd, ref = make_pair(shape, nTheta=360, nRho=180, nBands=12)
    pats = _pats(3, shape, dtype, seed=31, family="HCP")
    rad = bd.device_plan().radon(_dev(pats)).cpu().numpy()
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    assert np.array_equal(rad, want)
    got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(_dev(pats)).items()}
    res, st, C64, _ = oracle64(ref, pats)
    total, ties = compare_bands(got, res, C64, ref.padding, ref.peakPad)
    assert ties <= 1
    np.testing.assert_allclose(got['stats'][:, 2], st[2], rtol=1e-5)


@pytest.mark.parametrize("shape,dtype,n", [((240, 240), "uint16", 7), ((100, 120), "uint16", 5),
                                           ((61, 67), "uint8", 11), ((160, 160), "uint8", 3)])
def test_row_segments_identical(shape, dtype, n):
    """The integer gather over S row segments (partial sums added across S launches) gives
    the same float32 Radon as one launch over the whole pattern, with and without the fast
    background: uneven segment heights, a row length that defeats the 8-byte staging loads,
    ragged packs (n not a multiple of the pack size) and lines with no pixels."""
    import torch
    print(f"row segments {shape} {dtype} n={n}", flush=True)
    kw = dict(nTheta=360, nRho=180, nBands=12) if shape[0] == 240 else dict(nRho=130)
    bd, ref = make_pair(shape, **kw)
    pats = _pats(n, shape, dtype, seed=43)
    plan = bd.device_plan()
    x = _dev(pats)
    bg = np.linspace(0.0, 9.5, shape[0] * shape[1]).reshape(shape)
    for with_bg in (False, True):
        plan.set_background(bg if with_bg else None, False)
        outs = []
        for S in (1, 2, 3, 4, 8):
            plan.set_row_segments(S)
            outs.append(plan.radon(x).cpu().numpy())
        plan.set_row_segments(0)
        outs.append(plan.radon(x).cpu().numpy())
        for S, r in zip((2, 3, 4, 8, "auto"), outs[1:]):
            assert np.array_equal(r, outs[0]), (with_bg, S)
        if shape in ((100, 120), (160, 160)):
            assert (outs[0] == -1.0).any()   # lines with no pixels are written too
    plan.set_background(None, False)
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    plan.set_row_segments(3)
    assert np.array_equal(plan.radon(x).cpu().numpy(), want)
    plan.set_row_segments(0)


@pytest.mark.parametrize("shape,dtype,mode", [((240, 240), "float32", None), ((160, 160), "float32", None),
                                              ((100, 120), "uint16", "exact"), ((240, 240), "uint8", "exact")])
def test_float_row_segments_bit_identical(shape, dtype, mode):
    """Float sums (float32 patterns, the exact background mode) in row segments: each segment
    continues the float64 sum of the ones before it (a line's entries run in plan order,
    which is row-major), so the Radon equals the reference restatement's bit for bit for
    every segment count -- also where a whole 240x240 float pattern does not fit in LDS."""
    print(f"float row segments {shape} {dtype} {mode}", flush=True)
    kw = dict(nTheta=360, nRho=180, nBands=12) if shape[0] == 240 else {}
    bd, ref = make_pair(shape, background_mode=mode or "fast", **kw)
    pats = _pats(5, shape, dtype, seed=71)
    bg = None
    if mode:
        bg = np.linspace(-40.0, 55.5, shape[0] * shape[1]).reshape(shape) + 0.1234567
        bd.backgroundsub = bg
    plan = bd.device_plan()
    x, flags = bd._prepare(plan, _dev(pats))
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding, background=bg)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    for S in ((2, 3, 4) if shape[0] == 240 else (1, 2, 3)):
        plan.set_row_segments(S)
        assert np.array_equal(plan.radon(x, flags=flags).cpu().numpy(), want), S
    plan.set_row_segments(0)
    assert np.array_equal(plan.radon(x, flags=flags).cpu().numpy(), want)


@pytest.mark.parametrize("shape,dtype,kw,mode,segs", [
    ((120, 120), "uint16", {}, "fast", (0,)),
    ((120, 120), "uint8", dict(nTheta=190, nRho=95), None, (0, 2)),        # 18,050 cells: a partial unit
    ((240, 240), "uint16", dict(nTheta=360, nRho=180, nBands=12), "fast", (0, 2, 3)),
    ((240, 240), "uint16", dict(nTheta=360, nRho=180, nBands=12), "exact", (0,)),   # float row segments
    ((61, 67), "float32", {}, None, (0,)),
])
def test_line_units_identical(shape, dtype, kw, mode, segs):
    """The gather's line order in units of 1..64 neighbouring cells (the default 8) changes
    which lanes and batches carry which lines, never a sum: the Radon (whole patterns and row
    segments, integer and float sums, a cell count that leaves a partial unit) and the bands
    equal those of single-line units bit for bit; the plan reports each unit's padded slots."""
    print(f"line units {shape} {dtype} {kw} {mode} segments {segs}", flush=True)
    bd, ref = make_pair(shape, **kw)
    if mode:
        bd.background_mode = mode
        bd.backgroundsub = np.linspace(-3.0, 4.5, shape[0] * shape[1]).reshape(shape) + 0.25
    n = 6 if shape[0] == 240 else 21
    pats = _dev(_pats(n, shape, dtype, seed=61))
    plan = bd.device_plan()
    x, flags = bd._prepare(plan, pats)
    plan.set_line_unit(1)
    want = {S: (plan.set_row_segments(S), plan.radon(x, flags=flags).cpu().numpy())[1] for S in segs}
    plan.set_row_segments(0)
    bands = {k: v.cpu().numpy() for k, v in bd.find_bands_device(pats).items()}
    slots = {1: plan.info()["slots"]}
    for unit in (2, 8, 16, 64, 0):
        plan.set_line_unit(unit)
        info = plan.info()
        assert info["line_unit"] == (unit or 8) and info["slots"] >= slots[1]
        slots[unit] = info["slots"]
        for S in segs:
            plan.set_row_segments(S)
            assert np.array_equal(plan.radon(x, flags=flags).cpu().numpy(), want[S]), (unit, S)
        plan.set_row_segments(0)
        got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(pats).items()}
        for k in bands:
            assert np.array_equal(got[k], bands[k]), (unit, k)
    print(f"  entry slots per pattern by unit: {slots}", flush=True)
    from pyebsdindex_amd._hip import HipLibraryError
    for bad in (3, 128):
        with pytest.raises(HipLibraryError):
            plan.set_line_unit(bad)


def test_autotune_keeps_results():
    """The autotuner (pack size, work-group size, row segments, row pitch, peak tile) only
    changes launch shapes: the Radon and the bands after tuning equal those before."""
    from pyebsdindex_amd.autotune import autotune
    print("autotune 160x160 uint16", flush=True)
    shape = (160, 160)
    bd, ref = make_pair(shape)
    x = _dev(_pats(48, shape, "uint16", seed=47))
    plan = bd.device_plan()
    rad0 = plan.radon(x).cpu().numpy()
    b0 = {k: v.cpu().numpy() for k, v in bd.find_bands_device(x).items()}
    best = autotune(bd, x, sample=48, reps=1, gather_packs=(2, 4), gather_threads=(1024,),
                    row_segments=(1, 2, 3), peak_tiles=(45, 90), peak_threads=(512,))
    assert best["row_segments"] in (0, 1, 2, 3)
    assert np.array_equal(plan.radon(x).cpu().numpy(), rad0)
    b1 = {k: v.cpu().numpy() for k, v in bd.find_bands_device(x).items()}
    for k in b0:
        # the mean's float64 sum runs in another grouping under another peak tile or theta
        # block: the mean (and normmax, which divides by it) may differ in float32 rounding
        if k in ("stats", "normmax"):
            np.testing.assert_allclose(b0[k], b1[k], rtol=1e-6)
        else:
            assert np.array_equal(b0[k], b1[k]), k
    plan.set_row_segments(0)
    plan.set_stage_pitch(0)
    plan.set_peak_block(0)


@pytest.mark.parametrize("case", ["120-u16", "240-360x180", "uneven-theta", "kernel-5x5"])
def test_peak_blocks_agree(case):
    """Every theta block of the streamed peak kernel (wavefronts per pattern from the fewest
    to 16) gives the bands of the default block: identical positions, values, widths and
    min/max; the mean within float32 rounding of its float64 sum.  Out-of-range blocks and
    plans without the streamed kernel are refused."""
    from pyebsdindex_amd._hip import HipLibraryError
    print(f"peak blocks {case}", flush=True)
    kw, shape, dtype, n = {}, (120, 120), "uint16", 64
    if case == "240-360x180":
        shape, kw, n = (240, 240), dict(nTheta=360, nRho=180, nBands=12), 8
    elif case == "uneven-theta":
        kw = dict(nTheta=190, nRho=100)
    elif case == "kernel-5x5":
        kw = dict(tSigma=1.0, rSigma=1.0)
    bd, ref = make_pair(shape, **kw)
    plan = bd.device_plan()
    default, widest = plan.stream_block()
    rad = plan.radon(_dev(_pats(n, shape, dtype, seed=43)))
    want = {k: v.cpu().numpy() for k, v in plan.peaks(rad).items()}
    nT = plan.nT
    blocks = sorted({widest, default, -(-nT // 16), -(-nT // 7), 17, widest - 3} & set(range(-(-nT // 16), widest + 1)))
    for c in blocks:
        plan.set_peak_block(c)
        got = {k: v.cpu().numpy() for k, v in plan.peaks(rad).items()}
        for k in want:
            if k == "stats":
                assert np.array_equal(want[k][:, 0], got[k][:, 0]) and np.array_equal(want[k][:, 2], got[k][:, 2])
                np.testing.assert_allclose(want[k][:, 1], got[k][:, 1], rtol=2e-7, atol=1e-30)
            elif k == "normmax":
                np.testing.assert_allclose(want[k], got[k], rtol=1e-6, atol=1e-6)
            else:
                assert np.array_equal(want[k], got[k]), (c, k)
    for bad in (widest + 1, -(-nT // 16) - 1, -1):
        with pytest.raises(HipLibraryError):
            plan.set_peak_block(bad)
    plan.set_peak_block(0)
    assert {k: v.cpu().numpy() for k, v in plan.peaks(rad).items()}["max"].tolist() == want["max"].tolist()
    if case == "120-u16":            # a 13x9 kernel: only the tiled peak kernel applies
        bd13, _ = make_pair((60, 60), tSigma=2.0, rSigma=3.0)
        p13 = bd13.device_plan()
        assert p13.stream_block() is None
        with pytest.raises(HipLibraryError):
            p13.set_peak_block(20)


@pytest.mark.parametrize("case", ["120-u16", "60-u8-plateau", "240-360x180", "uneven-theta", "kernel-5x5"])
def test_peak_kernels_agree(case):
    """The streamed wavefront peak kernel and the theta-tiled work-group kernel compute the
    same stages with the same float32 arithmetic: identical bands, identical min/max and a
    mean that differs at most in float32 rounding of the float64 sum (summation order)."""
    import torch
    print(f"peak kernels {case}", flush=True)
    kw, shape, dtype, n = {}, (120, 120), "uint16", 64
    if case == "60-u8-plateau":
        shape, dtype = (60, 60), "uint8"
    elif case == "240-360x180":
        shape, kw, n = (240, 240), dict(nTheta=360, nRho=180, nBands=12), 8
    elif case == "uneven-theta":
        kw = dict(nTheta=190, nRho=100)
    elif case == "kernel-5x5":
        kw = dict(tSigma=1.0, rSigma=1.0)
    bd, ref = make_pair(shape, **kw)
    pats = _pats(n, shape, dtype, seed=41)
    if case == "60-u8-plateau":
        pats[: n // 2] = 77          # every masked cell ties: wavefront lists overflow and merge
        pats[n // 2: n // 2 + 2, :, 30:] = 200
    plan = bd.device_plan()
    rad = plan.radon(_dev(pats))
    outs = []
    for kind in (1, 2):
        plan.set_peak_kernel(kind)
        outs.append({k: v.cpu().numpy() for k, v in plan.peaks(rad).items()})
    plan.set_peak_kernel(0)
    a, b = outs
    for k in a:
        if k == "stats":
            assert np.array_equal(a[k][:, 0], b[k][:, 0]) and np.array_equal(a[k][:, 2], b[k][:, 2])
            np.testing.assert_allclose(a[k][:, 1], b[k][:, 1], rtol=2e-7, atol=1e-30)
        elif k == "normmax":
            np.testing.assert_allclose(a[k], b[k], rtol=1e-6, atol=1e-6)
        else:
            assert np.array_equal(a[k], b[k]), k
    torch.cuda.synchronize()


@pytest.mark.parametrize("kind", ["mask", "maskindex"])
def test_masked_plans(kind):
    """Plans built with a pattern mask (a circular detector: masked pixels leave the lines,
    radon_fast.py:167-193) or a mask index (pixels relabelled, here a flip of the pattern):
    the device Radon is bit-identical to the reference restatement with the same plan, and
    the bands agree with the FP64 oracle."""
    print(f"masked plan {kind}", flush=True)
    shape = (80, 80)
    H, W = shape
    if kind == "mask":
        yy, xx = np.mgrid[0:H, 0:W]
        kw = dict(patternmask=((yy - H / 2) ** 2 + (xx - W / 2) ** 2 < (0.45 * H) ** 2).astype(int))
    else:
        kw = dict(patternmaskindex=np.arange(H * W).reshape(H, W)[:, ::-1].copy())
    bd, ref = make_pair(shape, **kw)
    assert np.array_equal(bd.radonPlan.indexPlan, ref.radonPlan.indexPlan)
    pats = _pats(12, shape, "uint8", seed=31)
    rad = bd.device_plan().radon(_dev(pats)).cpu().numpy()
    p0, p1 = ref.padding
    want = ref.radonPlan.radon_faster(pats, ref.padding)[p0:-p0, p1:-p1].transpose(2, 0, 1)
    assert np.array_equal(rad, want)
    got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(_dev(pats)).items()}
    res, _, C64, _ = oracle64(ref, pats)
    total, ties = compare_bands(got, res, C64, ref.padding, ref.peakPad)
    assert ties <= max(1, total // 100)


def test_pass_sizes_identical():
    """find_bands gives the same bits whatever the pass size: the default (one pass while
    the Radons fit the workspace budget), equal passes from a workspace that holds 700 or
    1,000 Radons (2,500 patterns -> passes of 632/632/632/604 and 840/840/820), one pass
    per pattern, and the background modes."""
    import torch
    print("pass sizes 2500 x 120x120 uint16", flush=True)
    shape = (120, 120)
    bd, ref = make_pair(shape)
    pats = _pats(2500, shape, "uint16", seed=71)
    bg = pats[:256].astype(np.float64).mean(axis=0)
    x = _dev(pats)
    for mode in ("none", "fast", "exact"):
        bd.backgroundsub = None if mode == "none" else bg
        bd.background_mode = "exact" if mode == "exact" else "fast"
        plan = bd.device_plan()
        want = {k: v.cpu().numpy() for k, v in bd.find_bands_device(x).items()}
        for chunk in (700, 1000, 1):
            n = 7 if chunk == 1 else x.shape[0]
            got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(x[:n], chunk=chunk).items()}
            for k in want:
                assert np.array_equal(got[k], want[k][:n]), (mode, chunk, k)
        # a workspace that is not a whole number of Radons
        work = plan.workspace(x.shape[0], 900)
        work = work[: work.numel() - 7]
        got = {k: v.cpu().numpy() for k, v in bd.find_bands_device(x, work=work).items()}
        for k in want:
            assert np.array_equal(got[k], want[k]), (mode, "ragged workspace", k)
    torch.cuda.synchronize()
