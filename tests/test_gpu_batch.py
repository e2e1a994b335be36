"""Mixed-size batches: one persistent launch pair (and its hipGraph replay) must reproduce
each set's own find_bands bit for bit, and through it the FP64 oracle."""
import numpy as np
import pytest

from helpers import STD, compare_bands, make_pair, oracle64

pytestmark = pytest.mark.gpu

SETS = [((60, 60), "uint8"), ((80, 80), "uint16"), ((120, 120), "uint8"), ((160, 160), "uint16"),
        ((60, 60), "uint16"), ((100, 120), "uint8")]


def _sets(seed=40):
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    rng = np.random.default_rng(seed)
    dets, pats = [], []
    for k, (shape, dtype) in enumerate(SETS):
        pc = (rng.uniform(0.45, 0.55), rng.uniform(0.45, 0.55), rng.uniform(0.55, 0.7))
        n = int(rng.integers(3, 13))
        x = kikuchi_patterns(n, shape, dtype, seed=seed + k, family="BCC", pc=pc)
        dets.append(BandDetect(patDim=shape, **STD))
        pats.append(x.cuda())
    return dets, pats


def test_batch_equals_per_set_launches_and_graph_replays():
    import torch
    from pyebsdindex_amd.batch import BatchBandDetect
    dets, pats = _sets()
    want = [{k: v.cpu().numpy() for k, v in bd.find_bands_device(x).items()} for bd, x in zip(dets, pats)]
    batch = BatchBandDetect(dets, pats)
    batch.run()
    torch.cuda.synchronize()
    for i, (o, w) in enumerate(zip(batch.outs, want)):
        for k in w:
            assert np.array_equal(o[k].cpu().numpy(), w[k]), (i, k)
    # graph: captured on the first call, replayed; eager work and syncs in between
    for rep in range(3):
        for o in batch.outs:
            for v in o.values():
                v.zero_()
        torch.cuda.synchronize()
        batch.graph_launch()
        batch.graph_stream.synchronize()
        _ = dets[0].find_bands_device(pats[0])
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(batch.outs, want)):
            for k in w:
                assert np.array_equal(o[k].cpu().numpy(), w[k]), (rep, i, k)


def test_batch_against_fp64_oracle():
    dets, pats = _sets(seed=77)
    from pyebsdindex_amd.batch import BatchBandDetect
    import torch
    batch = BatchBandDetect(dets, pats)
    batch.run()
    torch.cuda.synchronize()
    for (shape, dtype), x, o in zip(SETS, pats, batch.outs):
        _, ref = make_pair(shape)
        res, st, C64, _ = oracle64(ref, x.cpu().numpy())
        got = {k: v.cpu().numpy() for k, v in o.items()}
        total, ties = compare_bands(got, res, C64, ref.padding, ref.peakPad)
        assert ties <= max(1, total // 50)


def test_batch_with_large_sets():
    """Sets of 1,024 patterns or more run the per-size gather and peak launches inside the
    batch (and its graph), row segments included (160x160 uint16); a small set stays in the
    merged launches.  Results equal each set's own
    find_bands bit for bit, eagerly and on graph replays with eager work in between."""
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd.batch import BatchBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print("batch with large sets", flush=True)
    spec = [((60, 60), "uint8", 2300), ((120, 120), "uint16", 2100), ((160, 160), "uint16", 2048),
            ((80, 80), "uint16", 1100), ((80, 80), "uint8", 40)]
    dets, pats = [], []
    for k, (shape, dtype, n) in enumerate(spec):
        pats.append(kikuchi_patterns(n, shape, dtype, seed=120 + k, family="BCC", device="cuda"))
        dets.append(BandDetect(patDim=shape, **STD))
    want = [{k: v.cpu().numpy() for k, v in bd.find_bands_device(x).items()} for bd, x in zip(dets, pats)]
    batch = BatchBandDetect(dets, pats)
    batch.run()
    torch.cuda.synchronize()
    for i, (o, w) in enumerate(zip(batch.outs, want)):
        for k in w:
            assert np.array_equal(o[k].cpu().numpy(), w[k]), (i, k)
    for rep in range(3):
        for o in batch.outs:
            for v in o.values():
                v.zero_()
        torch.cuda.synchronize()
        batch.graph_launch()
        batch.graph_stream.synchronize()
        _ = dets[0].find_bands_device(pats[0][:64])
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(batch.outs, want)):
            for k in w:
                assert np.array_equal(o[k].cpu().numpy(), w[k]), (rep, i, k)


def test_batch_pins_its_plans():
    """A batch holds device pointers into its plans: while it exists the plans refuse every
    set-up change, and a find_bands on one of them with a larger N (its row-segment partial
    sums grow; the outgrown buffer is kept until the plan goes) leaves the batch's eager run
    and graph replays bit-identical to per-set launches.  After the batch is destroyed the
    plans accept changes again."""
    import gc
    import torch
    from pyebsdindex_amd import BandDetect
    from pyebsdindex_amd._hip import HipLibraryError
    from pyebsdindex_amd.batch import BatchBandDetect
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print("batch pins its plans", flush=True)
    spec = [((240, 240), "uint16", 24), ((120, 120), "uint16", 2100), ((60, 60), "uint8", 30)]
    dets, pats = [], []
    for k, (shape, dtype, n) in enumerate(spec):
        kw = dict(STD, nTheta=360, nRho=180, nBands=12) if shape[0] == 240 else STD
        pats.append(kikuchi_patterns(n, shape, dtype, seed=150 + k, family="BCC", device="cuda"))
        dets.append(BandDetect(patDim=shape, **kw))
    want = [{k: v.cpu().numpy() for k, v in bd.find_bands_device(x).items()} for bd, x in zip(dets, pats)]
    batch = BatchBandDetect(dets, pats)
    plan = dets[0].device_plan()
    for call in (lambda: plan.set_row_segments(2), lambda: plan.set_tuning(gather_pack=2),
                 lambda: plan.set_stage_pitch(241), lambda: plan.set_peak_kernel(1),
                 lambda: plan.set_background(np.zeros((240, 240)), False)):
        with pytest.raises(HipLibraryError, match="bound to a batch"):
            call()
    # a larger find_bands on the 240x240 plan (row segments: its partial sums grow)
    big = kikuchi_patterns(600, (240, 240), "uint16", seed=160, family="BCC", device="cuda")
    ref_big = {k: v.cpu().numpy() for k, v in dets[0].find_bands_device(big).items()}
    for rep in range(3):
        for o in batch.outs:
            for v in o.values():
                v.zero_()
        # no explicit synchronisation: the replay's private stream is ordered both ways
        # with the current stream (the zeroing before it, the reads after it)
        batch.graph_launch() if rep else batch.run()
        for i, (o, w) in enumerate(zip(batch.outs, want)):
            for k in w:
                assert np.array_equal(o[k].cpu().numpy(), w[k]), (rep, i, k)
    again = {k: v.cpu().numpy() for k, v in dets[0].find_bands_device(big).items()}
    for k in ref_big:
        assert np.array_equal(again[k], ref_big[k]), k
    del batch
    gc.collect()
    plan.set_row_segments(2)
    plan.set_row_segments(0)
