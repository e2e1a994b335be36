"""NLPAR host logic without a device: the saturation thresholds handed to the kernels, the
row bands of a scan, the automask and the output file name (pyebsdindex_amd/nlpar.py)."""
import numpy as np
import pytest

from oracle import nlpar_ref
from pyebsdindex_amd import nlpar


def test_sigma_threshold_reproduces_float64_comparison():
    """sigma_numba compares float32 pixels with a float64 mxval (nlpar_cpu.py:763-767,
    Numba's float64 unification); the float32 threshold passed to the kernel must give the
    same answer for every float32 value, including those next to it."""
    rng = np.random.default_rng(0)
    for mx in list(rng.uniform(1, 70000, 200).astype(np.float32)) + [np.float32(255), np.float32(65535)]:
        for prot in (True, False):
            t64 = float(mx) * 0.9961 if prot else float(mx) + 1.0
            sat = np.float32(nlpar._sat_sigma(mx, prot))
            f = np.float32(t64)
            cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf)),
                     np.float32(np.floor(t64)), np.float32(np.ceil(t64))]
            for v in cands:
                assert (v < sat) == (float(v) < t64), (mx, prot, v)


def test_nlpar_threshold_is_float32_product():
    """nlpar_nb's mxval stays float32 (nlpar_cpu.py:866-870)."""
    for mx in (np.float32(255), np.float32(1234.5), np.float32(65535)):
        assert nlpar._sat_nlpar(mx, True) == float(mx * np.float32(0.999))
        assert nlpar._sat_nlpar(mx, False) == float(mx + np.float32(1.0))


def test_automask_matches_oracle():
    for h, w in ((60, 60), (61, 80), (120, 120)):
        assert np.array_equal(nlpar.makeautomask(h, w), nlpar_ref.makeautomask(h, w))


class _File:
    def __init__(self, pattern_bytes):
        self.pattern_bytes = pattern_bytes


@pytest.mark.parametrize("nrows,ncols,sr,budget_rows", [(12, 10, 3, 1), (40, 17, 3, 5), (9, 30, 4, 2),
                                                         (50, 20, 2, 50)])
def test_bands_hold_every_window(nrows, ncols, sr, budget_rows):
    """Row bands cover the scan once; the rows loaded for a band (nlpar.calcnlpar) hold
    every computed row's whole-scan window, at the same relative place, so a band run gives
    the whole-scan result."""
    nl = nlpar.NLPAR(nrows=nrows, ncols=ncols, band_bytes=(budget_rows + 2 * sr) * ncols * 100)
    bands = nl._bands(_File(100), sr)
    rows = [r for r0, r1 in bands for r in range(r0, r1)]
    assert rows == list(range(nrows))
    ws = 2 * sr + 1
    for r0, r1 in bands:
        a = max(0, min(r0 - sr, nrows - ws))
        b = min(nrows, max(r1 + sr, a + ws))
        assert b - a >= ws
        for j in range(r0, r1):
            g0 = max(j - sr, 0) - max(j + sr - (nrows - 1), 0)      # whole-scan window start
            jr = j - a
            l0 = max(jr - sr, 0) - max(jr + sr - (b - a - 1), 0)     # window start inside the band
            assert a + l0 == g0 and g0 >= a and g0 + ws <= b, (r0, r1, j)


def test_sigma_bands_overlap_two_rows():
    """calcsigma's bands add 2 nn rows each side, so the neighbours of a kept point have
    their whole +-nn neighbourhood in the band."""
    nrows, nn = 23, 1
    nl = nlpar.NLPAR(nrows=nrows, ncols=8, band_bytes=(3 + 4 * nn) * 8 * 50)
    for r0, r1 in nl._bands(_File(50), 2 * nn):
        a, b = max(0, r0 - 2 * nn), min(nrows, r1 + 2 * nn)
        for j in range(r0, r1):
            for jj in range(max(0, j - nn), min(nrows, j + nn + 1)):
                assert max(0, jj - nn) >= a and min(nrows, jj + nn + 1) <= b


def test_output_name(tmp_path):
    from pyebsdindex_amd.ebsd_pattern import UPFile, get_pattern_file_obj
    fn = tmp_path / "scan.up1"
    w = UPFile(fn)
    w.patternW, w.patternH, w.nCols, w.nRows, w.xStep, w.yStep = 8, 8, 3, 2, 1.0, 1.0
    w.write_header()
    w.pat_writer(np.zeros((6, 8, 8), np.uint8), 0, 6, np.uint8)
    nl = nlpar.NLPAR(str(fn), lam=0.75, searchradius=2)
    f = nl.getinfileobj()
    nl.setoutfile(f)
    assert str(nl.filepathout).endswith("scan_NLPAR_l0.75sr2.up1")
    out = get_pattern_file_obj(str(nl.filepathout))
    assert (out.nCols, out.nRows, out.patternW, out.patternH) == (3, 2, 8, 8)
    assert out.nPatterns == 6
    with pytest.raises(ValueError):
        nl.setoutfile(f, filepath=str(fn))
