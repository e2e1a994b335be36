"""The band-detection plan object as PyEBSDIndex's indexers and schedulers use it (CPU only).

The reference checks its OpenCL plan only by subclassing (pyebsdindex/tests/test_package.py:
32-36).  This package's ``BandDetect`` is not a subclass, so the contract is checked here
directly: every attribute and keyword the indexer, the Ray scheduler, the gnomonic
correction and ``pcopt`` read from ``bandDetectPlan`` (pyebsdindex/_ebsd_index_single.py:
245-251,398-413,462,550,697-720; _ebsd_index_parallel.py:294-313,356,403,432,768-830;
gnomonic_correction.py:71,94; pcopt.py:94,114,192), the device descriptor
``getopenclparam`` returns (band_detect_cl.py:719-722, opencl/openclparam.py), and pickling
(Ray's object store pickles the indexer with its plan, _ebsd_index_parallel.py:396).
"""
import inspect
import pickle
import threading
import types

import numpy as np
import pytest

from pyebsdindex_amd import BandDetect
from pyebsdindex_amd import band_detect as bdmod
from pyebsdindex_amd.radon_fast import Radon

# attributes read through ``indexer.bandDetectPlan.<name>`` anywhere in the reference package
PLAN_ATTRIBUTES = ("backgroundsub", "dataType", "kernel", "nBands", "nRho", "nTheta", "padding", "patDim",
                   "rSigma", "radonPlan", "rdnmask", "rhoMaskFrac", "rhoMax", "useCPU")
# the fields of the band record the indexer reads or writes (_ebsd_index_single.py)
RECORD_FIELDS = ("id", "max", "normmax", "maxloc", "avemax", "aveloc", "pqmax", "width", "theta", "rho",
                 "valid", "band_match_index", "bandfit")


def _plan(**kw):
    args = dict(patDim=(120, 120), nTheta=180, nRho=90, tSigma=1.0, rSigma=2.0, rhoMaskFrac=0.1, nBands=9)
    args.update(kw)
    return BandDetect(**args)


def test_plan_exposes_what_the_indexer_reads():
    bd = _plan()
    for name in PLAN_ATTRIBUTES:
        assert hasattr(bd, name), name
    assert tuple(np.asarray(bd.patDim)) == (120, 120)
    assert bd.nTheta == 180 and bd.nRho == 90 and bd.nBands == 9
    assert bd.useCPU is False          # _ebsd_index_parallel.py:310-311 reads it as a GPU switch
    assert bd.backgroundsub is None
    assert np.asarray(bd.padding).shape == (2,)
    assert bd.kernel.ndim == 3 and bd.kernel.dtype == np.float32
    assert bd.rdnmask.shape == (90, 180)
    assert bd.rhoMax == pytest.approx(60.0)
    assert bd.dataType.names == RECORD_FIELDS
    assert bd.dataType["band_match_index"].shape == (bd.nPhases,)
    # the chunk estimate of _ebsd_index_parallel.py:801-805 uses these as numbers
    rdndim = np.array([bd.nTheta + 2 * bd.padding[1], bd.nRho + 2 * bd.padding[0]])
    memperpat = 4.0 * float(bd.patDim[0] * bd.patDim[1] + 6.0 * rdndim[0] * rdndim[1])
    assert memperpat > 0
    # the structured array the indexer allocates for the results (:432-434)
    out = np.zeros((4, bd.nBands), dtype=bd.dataType)
    assert out["maxloc"].shape == (4, 9, 2)
    # the indexer assigns a background array directly (:245-246)
    bd.backgroundsub = np.zeros((120, 120), np.float32)


def test_plan_methods_take_the_indexers_keywords():
    bd = _plan()
    # find_bands(pats, verbose=, chunksize=, clparams=, gpu_id=)  (_ebsd_index_single.py:705-707)
    p = inspect.signature(bd.find_bands).parameters
    for k in ("verbose", "chunksize", "clparams", "gpu_id"):
        assert k in p, k
    # band_detect_setup(patDim=, patternmask=, patternmaskindex=) and (patterns=)  (:401-409,550)
    p = inspect.signature(bd.band_detect_setup).parameters
    for k in ("patterns", "patDim", "patternmask", "patternmaskindex"):
        assert k in p, k
    # collect_background(fileobj=, patsIn=, nsample=)  (:249-251)
    p = inspect.signature(bd.collect_background).parameters
    for k in ("fileobj", "patsIn", "nsample"):
        assert k in p, k
    # radonPlan.radon2pole(banddata, PC=, vendor=)  (:716-718, pcopt.py:94,114)
    p = inspect.signature(bd.radonPlan.radon2pole).parameters
    for k in ("PC", "vendor"):
        assert k in p, k
    # GnomonicCorrection reads radonPlan.imDim (gnomonic_correction.py:71); the indexer's
    # save/restore reads the pattern mask and mask index (_ebsd_index_single.py:964-965)
    assert tuple(bd.radonPlan.imDim) == (120, 120)
    assert np.asarray(bd.radonPlan.mask).shape == (120, 120)
    assert np.array_equal(np.asarray(bd.radonPlan.maskindex).ravel(), np.arange(120 * 120))
    # a second set-up with new dimensions, as the indexer does for a new file (:697-700).  The
    # plan keeps the previous pattern mask and mask index, as the reference's does: smaller
    # patterns without new ones are refused by the mask-index check (radon_fast.py:189-190)
    with pytest.raises(Exception, match="max of index array"):
        _plan().band_detect_setup(patDim=(60, 80))
    bd.band_detect_setup(patDim=(60, 80), patternmask=np.ones((60, 80)),
                         patternmaskindex=np.arange(60 * 80).reshape(60, 80))
    assert tuple(np.asarray(bd.patDim)) == (60, 80) and tuple(bd.radonPlan.imDim) == (60, 80)
    assert bd.radonPlan.indexPlan.shape[:2] == (90, 180) and bd.rdnmask.shape == (90, 180)


def test_getopenclparam_without_gpu_is_none():
    """The scheduler takes None as 'no GPU' (_ebsd_index_parallel.py:291-296, :773)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert bdmod.getopenclparam() is None


def _fake_cuda(monkeypatch, mems=(288 * 2 ** 30, 256 * 2 ** 30)):
    import torch
    props = [types.SimpleNamespace(name=f"AMD Instinct MI355X #{i}", total_memory=m, multi_processor_count=256)
             for i, m in enumerate(mems)]
    monkeypatch.setattr(torch.cuda, "is_available", lambda: True)
    monkeypatch.setattr(torch.cuda, "device_count", lambda: len(props))
    monkeypatch.setattr(torch.cuda, "get_device_properties", lambda i: props[int(i)])
    return props


def test_device_descriptor_serves_the_scheduler(monkeypatch):
    """The reads of _ebsd_index_parallel.py:294-313 and __optimizegpuchunk__ (:768-830) on
    the descriptor, with two devices of different memory sizes."""
    props = _fake_cuda(monkeypatch)
    clparam = bdmod.getopenclparam()
    assert isinstance(clparam, bdmod.HipParam)
    # :301-306 -- device count and ids
    assert len(clparam.gpu) == 2 and clparam.ngpu == 2
    gpu_id = np.arange(len(clparam.gpu), dtype=int)
    assert ",".join(str(g) for g in gpu_id) == "0,1"
    # :779-786 -- iteration and indexing by numpy integers
    assert [g.global_mem_size for g in clparam.gpu] == [p.total_memory for p in props]
    assert clparam.gpu[np.int64(1)].global_mem_size == props[1].total_memory
    gmem = min(g.global_mem_size for g in (clparam.gpu[g] for g in np.atleast_1d(gpu_id)))
    assert gmem == 256 * 2 ** 30
    # :824 -- integrated-GPU switch
    assert clparam.gpusharedmem is False
    # the worker's call (:853, band_detect_cl.py:61-63): a factory, then get_queue(gpu_id=)
    q = bdmod.getopenclparam().get_queue(gpu_id=1)
    assert q.gpu_id == 1
    assert bdmod.HipParam(gpu_id=1).gpu_id == 1


def test_find_bands_takes_device_from_clparams(monkeypatch):
    """find_bands(clparams=HipParam(gpu_id=k)) selects device k, as the OpenCL plan takes
    its queue from clparams (band_detect_cl.py:61-63); an explicit gpu_id wins."""
    _fake_cuda(monkeypatch)
    bd = _plan()
    seen = []

    def fake_plan(gpu_id=None):
        seen.append(gpu_id)
        raise RuntimeError("stop")
    monkeypatch.setattr(bd, "device_plan", fake_plan)
    pats = np.zeros((2, 120, 120), np.uint8)
    for kw in (dict(clparams=bdmod.HipParam(gpu_id=1)), dict(clparams=bdmod.HipParam(gpu_id=1), gpu_id=0),
               dict()):
        with pytest.raises(RuntimeError, match="stop"):
            bd.find_bands(pats, **kw)
    assert seen == [1, 0, None]


def test_plan_pickles_without_device_state():
    bd = _plan()
    bd.backgroundsub = np.full((120, 120), 3.0, np.float32)
    # stand-ins for the per-process state a used plan carries (library handles, pinned memory)
    bd._device_plans[0] = threading.Lock()
    bd._pipelines = {"k": threading.Lock()}
    bd.radonPlan._device_plans = {0: threading.Lock()}
    bd2 = pickle.loads(pickle.dumps(bd))
    assert bd2._device_plans == {} and "_pipelines" not in bd2.__dict__
    assert "_device_plans" not in bd2.radonPlan.__dict__
    for name in PLAN_ATTRIBUTES:
        a, b = getattr(bd, name), getattr(bd2, name)
        if isinstance(a, np.ndarray):
            assert np.array_equal(a, b), name
        elif name != "radonPlan":
            assert a == b, name
    assert np.array_equal(bd.radonPlan.indexPlan, bd2.radonPlan.indexPlan)
    # the original keeps its state
    assert 0 in bd._device_plans and "_pipelines" in bd.__dict__


def test_new_setup_drops_device_state():
    """Device plans and host pipelines are per geometry: a set-up that changes the geometry,
    the kernel or the band count forgets them; one that changes nothing keeps them."""
    bd = _plan()
    smaller = dict(patDim=(100, 120), patternmask=np.ones((100, 120)),
                   patternmaskindex=np.arange(100 * 120).reshape(100, 120))
    for change, drops in ((dict(patDim=(120, 120)), False), (smaller, True), (dict(patDim=(100, 120)), False),
                          (dict(rSigma=2.5), True), (dict(nBands=9), False), (dict(nBands=8), True)):
        bd._device_plans[0] = "plan"
        bd._pipelines = {"k": "pipe"}
        bd.band_detect_setup(**change)
        assert (bd._device_plans == {}) == drops, change
        assert ("_pipelines" not in bd.__dict__) == drops, change
    r = Radon(imageDim=(60, 60), nTheta=180, nRho=90)
    r._device_plans = {0: "plan"}
    r.radon_plan_setup(nTheta=90)
    assert "_device_plans" not in r.__dict__


@pytest.mark.parametrize("libtype", ["FCC", "BCC", "HCP"])
def test_phase_exposes_what_the_indexer_reads(libtype):
    """The phase objects of ``EBSDIndexer.phaseLib``: the call of _ebsd_index_single.py:791-795,
    the pole lookups of getmatchedpole (:659,668) and the phase header of ebsdfile.py:24-37."""
    from pyebsdindex_amd import tripletvote as tv
    ph = tv.addphase(libtype=libtype)
    p = inspect.signature(ph.bandindex).parameters
    for k in ("band_intensity", "band_widths", "verbose"):
        assert k in p, k
    for name in ("phasename", "lauecode", "latticeparameter", "polefamilies", "npolefamilies", "spacegroup",
                 "angTol", "nband_earlyexit", "completelib"):
        assert hasattr(ph, name), name
    assert ph.phasename == libtype
    lib = ph.completelib
    assert lib["poles"].shape[0] == lib["polesCart"].shape[0] and lib["polesCart"].shape[1] == 3
    assert np.asarray(ph.polefamilies).shape[0] == ph.npolefamilies
    assert np.asarray(ph.latticeparameter).shape == (6,)
    # pickles without its device copies
    ph._dev[("complete", "cuda:0")] = threading.Lock()
    ph2 = pickle.loads(pickle.dumps(ph))
    assert ph2._dev == {} and ph2.lauecode == ph.lauecode
    assert np.array_equal(ph2.angtriplets["angles"], ph.angtriplets["angles"])


def test_auto_pass_size_rule():
    """The pass size taken when a caller names no chunk: all patterns while their Radons fit
    the workspace budget, else whole gather rounds (4 patterns per CU) that fit it."""
    cells = 180 * 90                 # 64,800-byte float32 Radon
    gib = 1 << 30
    # the headline workload: 16,384 Radons are 1.06e9 bytes < 1 GiB -> one pass
    assert bdmod.auto_pass_size(16384, cells, 256, gib) == 16384
    assert bdmod.auto_pass_size(1, cells, 256, gib) == 1
    assert bdmod.auto_pass_size(0, cells, 256, gib) == 1
    # 20,000 do not fit: 1 GiB holds 16,570 -> 16 rounds of 1,024
    assert bdmod.auto_pass_size(20000, cells, 256, gib) == 16384
    # 240x240 with Radon 360x180 (259,200 bytes): 4,142 fit -> 4 rounds
    assert bdmod.auto_pass_size(8192, 360 * 180, 256, gib) == 4096
    # a budget below one round keeps what fits
    assert bdmod.auto_pass_size(5000, cells, 256, 100 * cells * 4) == 100
    # every answer fits the budget and never exceeds N
    for n in (1, 7, 1023, 1024, 4097, 50000):
        for b in (cells * 4, 10 ** 6, 10 ** 8, gib):
            k = bdmod.auto_pass_size(n, cells, 256, b)
            assert 1 <= k <= max(1, n) and (k * cells * 4 <= b or k == 1)


def test_record_layout_of_the_structured_array():
    """The drop-in call has the device write its records into the returned array only for a
    dataType with the reference's packed layout (band_detect.py:110-115): record_phases
    gives its phase count, and None for a layout that differs in order, offsets or types."""
    for nph in (1, 2, 10):
        bd = BandDetect(patDim=(60, 60), nTheta=180, nRho=90, nPhases=nph)
        assert bd.dataType.itemsize == 49 + 16 * nph
        assert bdmod.record_phases(bd.dataType) == nph
    dt = BandDetect(patDim=(60, 60), nTheta=180, nRho=90).dataType
    aligned = np.dtype([(n, dt.fields[n][0]) for n in dt.names], align=True)
    assert bdmod.record_phases(aligned) is None
    reordered = np.dtype([(n, dt.fields[n][0]) for n in reversed(dt.names)])
    assert bdmod.record_phases(reordered) is None
    wider = np.dtype([(n, np.float64 if n == "theta" else dt.fields[n][0]) for n in dt.names])
    assert bdmod.record_phases(wider) is None
    assert bdmod.record_phases(np.dtype(np.float32)) is None
