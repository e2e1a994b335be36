"""The C-ABI library loads and exports every entry point declared in include/*.h; host-side
set-up mirrors the reference; no CPU fallback exists.  No GPU compute here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from helpers import STD
from oracle import ref_cpu


def _declared(header="*.h"):
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", header)):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(ebsd_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from pyebsdindex_amd import _hip
    lib = _hip.load()
    declared = _declared("ebsd_radon.h")
    assert len(declared) >= 11
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_hip.SYMBOLS) <= set(declared)
    nl = _hip.load_nlpar()
    declared_nl = _declared("ebsd_nlpar.h")
    assert set(declared_nl) == set(_hip.NLPAR_SYMBOLS)
    for name in declared_nl:
        assert hasattr(nl, name), name
    tv = _hip.load_tripvote()
    declared_tv = _declared("ebsd_tripvote.h")
    assert set(declared_tv) == set(_hip.TRIPVOTE_SYMBOLS)
    for name in declared_tv:
        assert hasattr(tv, name), name
    # the measurement-only probe (LDS read and VALU issue ceilings for the bench's rooflines)
    from pyebsdindex_amd.build import PROBE_LIB
    pr = ctypes.CDLL(PROBE_LIB)
    declared_pr = _declared("ebsd_probe.h")
    assert declared_pr == ["ebsd_probe_lds_read", "ebsd_probe_valu_issue"]
    for name in declared_pr:
        assert hasattr(pr, name), name
    assert pr.ebsd_probe_lds_read(3, 1, 1, None, None) == -1   # bad arguments, no launch
    assert pr.ebsd_probe_valu_issue(4, 1, 1, None, None) == -1
    # every header's declarations are covered by one of the libraries
    assert set(_declared()) == set(declared) | set(declared_nl) | set(declared_tv) | set(declared_pr)


def test_tripvote_error_paths_without_device_compute():
    from pyebsdindex_amd import _hip
    tv = _hip.load_tripvote()
    args = lambda npats, nb, ntrip, nfam: (None, None, npats, nb, None, None, ntrip, nfam, 2.0,
                                           None, None, 0, 0.0, None, None, None, None, None, None)
    assert tv.ebsd_tripvote(*args(4, 65, 10, 4)) < 0
    assert b"nbands" in tv.ebsd_tripvote_last_error()
    assert tv.ebsd_tripvote(*args(4, 9, 10, 1)) < 0
    assert b"nfam" in tv.ebsd_tripvote_last_error()
    assert tv.ebsd_tripvote(*args(4, 9, 10, 4)) < 0
    assert b"null" in tv.ebsd_tripvote_last_error()
    assert tv.ebsd_tripvote(*args(4, 64, 10, 200)) < 0
    assert b"LDS" in tv.ebsd_tripvote_last_error()
    assert tv.ebsd_tripvote(*args(0, 9, 10, 4)) == 0          # no patterns: nothing to do


def test_nlpar_error_paths_without_device_compute():
    from pyebsdindex_amd import _hip
    nl = _hip.load_nlpar()
    assert nl.ebsd_nlpar_sigma(None, 0, 3, 3, 16, None, 16, 1, 1.0, None, None, None, None, 0, None) < 0
    assert b"null" in nl.ebsd_nlpar_last_error()
    # window larger than the chunk
    assert nl.ebsd_nlpar(None, 0, 5, 5, 16, None, 16, None, 3, 0.7, 0.0, 0.0, 1.0, 0, 0, 5, 5, None, 0,
                         None, None, 0, None) < 0
    assert b"2*searchradius+1" in nl.ebsd_nlpar_last_error()
    assert nl.ebsd_nlpar_max(None, 0, 1, 16, None, None) < 0
    assert nl.ebsd_nlpar_workspace_bytes(10, 3600) == 3840 + 256 + 256 + 2560   # lists of 64 int32


def test_error_paths_without_device_compute():
    from pyebsdindex_amd import _hip
    lib = _hip.load()
    out = ctypes.c_void_p()
    rc = lib.ebsd_plan_create(60, 60, 180, 90, None, 60, None, None, 9, 5, 19, 11, 19, 11, 9, None, None,
                              ctypes.byref(out))
    assert rc < 0 and b"null" in lib.ebsd_last_error()
    assert lib.ebsd_find_bands(None, None, 0, 1, None, None, 0, None) < 0
    assert lib.ebsd_radon(None, None, 0, 1, None, None) < 0
    assert lib.ebsd_workspace_bytes(None, 10) == 0


@pytest.mark.parametrize("shape,kw", [((60, 60), {}), ((100, 120), {}), ((120, 120), dict(rhoMaskFrac=0.15)),
                                      ((80, 80), dict(rhoMaskFrac=-0.2)), ((60, 60), dict(rhoMaskFrac=2.0)),
                                      ((60, 60), dict(tSigma=None, rSigma=None))])
def test_setup_matches_reference_restatement(shape, kw):
    from pyebsdindex_amd import BandDetect
    p = dict(STD)
    p.update(kw)
    bd = BandDetect(patDim=shape, **p)
    ref = ref_cpu.RefBandDetect(shape, **p)
    assert np.array_equal(bd.padding, ref.padding) and np.array_equal(bd.peakPad, ref.peakPad)
    assert np.array_equal(bd.kernel, ref.kernel)
    assert np.array_equal(bd.rdnmask, ref.rdnmask)
    assert bd.rhoMax == ref.rhoMax and bd.dRho == ref.dRho
    assert np.array_equal(bd.radonPlan.theta, ref.radonPlan.theta)
    assert np.array_equal(bd.radonPlan.rho, ref.radonPlan.rho)
    assert bd.dataType == ref.dataType


def test_setup_with_pattern_mask():
    from pyebsdindex_amd import BandDetect
    H, W = 60, 60
    yy, xx = np.mgrid[0:H, 0:W]
    mask = ((yy - H / 2) ** 2 + (xx - W / 2) ** 2 < (0.45 * H) ** 2).astype(int)
    bd = BandDetect(patDim=(H, W), patternmask=mask, **STD)
    ref = ref_cpu.RefBandDetect((H, W), patternmask=mask, **STD)
    assert np.array_equal(bd.radonPlan.indexPlan, ref.radonPlan.indexPlan)
    assert np.array_equal(bd.rdnmask, ref.rdnmask)


def test_no_cpu_fallback_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd._hip import HipLibraryError
    bd = BandDetect(patDim=(60, 60), **STD)
    with pytest.raises(HipLibraryError):
        bd.find_bands(np.zeros((2, 60, 60), dtype=np.uint8))


def test_missing_library_fails_loudly(tmp_path):
    from pyebsdindex_amd import _hip
    with pytest.raises(_hip.HipLibraryError):
        _hip.load.__wrapped__(str(tmp_path / "nope.so")) if hasattr(_hip.load, "__wrapped__") else \
            _load_fresh(_hip, str(tmp_path / "nope.so"))


def _load_fresh(_hip, path):
    saved = _hip._lib
    _hip._lib = None
    try:
        return _hip.load(path)
    finally:
        _hip._lib = saved


def test_background_fit_recovers_gaussian_surface():
    """backsub_fit (band_detect.py:350-408) on an exact Gaussian-on-a-plane returns it, zero-mean."""
    from pyebsdindex_amd import BandDetect
    ny, nx = 60, 70
    yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
    surf = 300 * np.exp(-((xx - 33) ** 2) / (2 * 20.0 ** 2) - ((yy - 28) ** 2) / (2 * 25.0 ** 2)) + 50 + 0.3 * xx - 0.2 * yy
    bd = BandDetect(patDim=(ny, nx), **STD)
    fit = bd.backsub_fit(surf)
    assert abs(fit.mean()) < 1e-9
    # the fit runs on the 3x3 median-filtered mean pattern, so it is exact only up to the
    # median filter's deviation from a curved surface (~1e-3 of the range here)
    assert np.abs(fit - (surf - surf.mean())).max() < 1e-3 * np.ptp(surf)
