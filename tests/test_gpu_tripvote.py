"""Device triplet vote (libebsd_tripvote.so through pyebsdindex_amd.tripletvote) against
the oracle's C restatement of BandIndexer._tripvote_numba (tripletvote.py:1019-1207).

Bar: the vote counts and the family of each band are identical; the weighted votes,
cm and rank are identical to within 1e-4 absolute (float32 sums of weights near angTol;
the device adds in the reference's order, so almost every entry is bit-identical -- at
least 99.9 % of them must be).  The only source of difference is the float64 arccos of
the device and host math libraries, rounded to float32."""
import numpy as np
import pytest

from helpers import cubic_bands

pytestmark = pytest.mark.gpu


def _cases():
    return [("FCC", 9, 0.2, 0.0), ("BCC", 9, 0.3, 0.05), ("FCC", 12, 0.3, 0.05), ("BCC", 3, 0.2, 0.0),
            ("FCC", 20, 0.5, 0.1)]


def _compare(ind, bn, inten, tol=None):
    import torch
    from oracle import tripvote_ref
    dev = torch.device("cuda", 0)
    tol = ind.angTol if tol is None else tol
    bnd = torch.as_tensor(bn, device=dev)
    intd = None if inten is None else torch.as_tensor(inten, device=dev)
    from pyebsdindex_amd.tripletvote import tripvote_device
    got = [t.cpu().numpy() for t in tripvote_device(bnd, intd, ind.lut, tol, ind.angtriplets["angles"],
                                                       ind.angtriplets["familyid"], ind.npolefamilies)]
    want = tripvote_ref.tripvote(bn, inten, ind.lut, tol, ind.angtriplets["angles"],
                                 ind.angtriplets["familyid"], ind.npolefamilies)
    accW, fam, rank, cm, acc = got
    assert np.array_equal(acc, want[4])
    assert np.array_equal(fam, want[1])
    for g, w in ((accW, want[0]), (rank, want[2]), (cm, want[3])):
        assert np.abs(g - w).max() <= 1e-4
        assert np.mean(g == w) >= 0.999
    return got, want


@pytest.mark.parametrize("phase,nb,noise,invalid", _cases())
def test_vote_matches_oracle(phase, nb, noise, invalid):
    from pyebsdindex_amd import tripletvote as tv
    ind = tv.addphase(phase)
    npats = 300
    print(f"tripvote case {phase} nb={nb} noise={noise} invalid={invalid}", flush=True)
    bn, _ = cubic_bands(ind, npats, nb, seed=nb * 7 + int(noise * 10), noise_deg=noise)
    rng = np.random.default_rng(nb)
    inten = rng.uniform(0.1, 1.0, (npats, nb)).astype(np.float32)
    inten[rng.random((npats, nb)) < invalid] = 0.0
    _compare(ind, bn, inten)


def test_vote_defaults_and_single_pattern():
    import torch
    from pyebsdindex_amd import tripletvote as tv
    ind = tv.addphase("FCC")
    bn, fam_true = cubic_bands(ind, 1, 9, seed=2)
    _compare(ind, bn[0], None)
    # BandIndexer.tripvote: the library uploaded once, the same outputs
    dev = torch.device("cuda", 0)
    a = ind.tripvote(torch.as_tensor(bn, device=dev))
    b = ind.tripvote(torch.as_tensor(bn, device=dev))
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.parametrize("tol", [0.5, 1.0, 3.0])
def test_other_tolerances(tol):
    from pyebsdindex_amd import tripletvote as tv
    ind = tv.addphase("BCC")
    bn, _ = cubic_bands(ind, 200, 9, seed=31, noise_deg=0.2)
    _compare(ind, bn, None, tol=tol)


def test_edge_shapes():
    """One band (no triplet), an empty library, 64 bands (the largest), two families."""
    import torch
    from pyebsdindex_amd import tripletvote as tv
    from oracle import tripvote_ref
    dev = torch.device("cuda", 0)
    ind = tv.addphase("FCC")
    bn, _ = cubic_bands(ind, 20, 1, seed=1)
    _compare(ind, bn, None)
    bn20, _ = cubic_bands(ind, 8, 20, seed=4, noise_deg=0.1)
    extra = np.random.default_rng(0).normal(size=(8, 44, 3))
    extra /= np.linalg.norm(extra, axis=2, keepdims=True)
    _compare(ind, np.ascontiguousarray(np.concatenate([bn20, extra.astype(np.float32)], axis=1)), None)
    bn9, _ = cubic_bands(ind, 10, 9, seed=6)
    empty_a, empty_i = np.zeros((0, 3)), np.zeros((0, 3), np.int64)
    got = tv.tripvote_device(torch.as_tensor(bn9, device=dev), None, ind.lut, 2.0, empty_a, empty_i, 4)
    want = tripvote_ref.tripvote(bn9, None, ind.lut, 2.0, empty_a, empty_i, 4)
    for g, w in zip(got, want):
        assert np.array_equal(g.cpu().numpy(), w)
    # a two-family library: only the {002} and {111} rows of FCC
    keep = np.all(ind.angtriplets["familyid"] < 2, axis=1)
    a2, i2 = ind.angtriplets["angles"][keep], ind.angtriplets["familyid"][keep]
    got = tv.tripvote_device(torch.as_tensor(bn9, device=dev), None, ind.lut, 2.0, a2, i2, 2)
    want = tripvote_ref.tripvote(bn9, None, ind.lut, 2.0, a2, i2, 2)
    assert np.array_equal(got[4].cpu().numpy(), want[4])
    assert np.abs(got[0].cpu().numpy() - want[0]).max() <= 1e-4


def test_large_batch_is_the_union_of_its_parts():
    """65,536 patterns in one launch give the same result as 16 launches of 4,096, and a
    sample of them equals the oracle."""
    import torch
    from pyebsdindex_amd import tripletvote as tv
    from oracle import tripvote_ref
    dev = torch.device("cuda", 0)
    ind = tv.addphase("FCC")
    base, _ = cubic_bands(ind, 512, 9, seed=77, noise_deg=0.3)
    rng = np.random.default_rng(5)
    bn = base[rng.integers(0, 512, 65536)]
    bnd = torch.as_tensor(bn, device=dev)
    whole = ind.tripvote(bnd)
    parts = [ind.tripvote(bnd[s:s + 4096]) for s in range(0, 65536, 4096)]
    for k in range(5):
        assert torch.equal(whole[k], torch.cat([p[k] for p in parts]))
    idx = rng.choice(65536, 256, replace=False)
    want = tripvote_ref.tripvote(bn[idx], None, ind.lut, 2.0, ind.angtriplets["angles"],
                                 ind.angtriplets["familyid"], ind.npolefamilies)
    assert np.array_equal(whole[4][idx].cpu().numpy(), want[4])
    assert np.array_equal(whole[1][idx].cpu().numpy(), want[1])


def test_bad_inputs_raise():
    import torch
    from pyebsdindex_amd import tripletvote as tv
    from pyebsdindex_amd._hip import HipLibraryError
    dev = torch.device("cuda", 0)
    ind = tv.addphase("FCC")
    bn = torch.zeros((2, 9, 3), dtype=torch.float32, device=dev)
    with pytest.raises(TypeError):
        tv.tripvote_device(bn.double(), None, ind.lut, 2.0, ind.angtriplets["angles"], ind.angtriplets["familyid"], 4)
    with pytest.raises(ValueError):
        tv.tripvote_device(bn, None, np.zeros((3, 3, 3, 3)), 2.0, ind.angtriplets["angles"],
                           ind.angtriplets["familyid"], 4)
    with pytest.raises(ValueError):            # ids 0..3 with nfam = 2, caught on the host
        tv.tripvote_device(bn, None, ind.lut, 2.0, ind.angtriplets["angles"], ind.angtriplets["familyid"], 2)
    with pytest.raises(HipLibraryError):       # one family: rejected by the library
        tv.tripvote_device(bn, None, ind.lut, 2.0, np.zeros((0, 3)), np.zeros((0, 3), np.int64), 1)


@pytest.mark.parametrize("extra", [[[0, 2, 4], [1, 3, 3]],
                                   [[0, 2, 4], [2, 2, 4], [1, 3, 3], [1, 3, 5], [1, 1, 5], [1, 2, 3]]])
def test_more_families(extra):
    """FCC plus further reflectors: 6 families (1,438 triplets, the 8-family register
    columns, library in LDS) and 10 families (the 16-family columns; the library is too
    large for LDS and is read from global memory)."""
    from pyebsdindex_amd import tripletvote as tv
    ind = tv.BandIndexer("FCC+", polefamilies=np.concatenate([tv.FCC_POLES, np.array(extra)]))
    assert ind.npolefamilies == 4 + len(extra)
    print(f"tripvote families {ind.npolefamilies}, {ind.angtriplets['angles'].shape[0]} triplets", flush=True)
    bn, _ = cubic_bands(ind, 120, 9, seed=len(extra), noise_deg=0.2)
    _compare(ind, bn, None)


@pytest.mark.parametrize("phase,tol", [("FCC", 2.0), ("BCC", 0.5), ("BCC", 3.0)])
def test_row_index_gives_identical_outputs(phase, tol):
    """The optional row index only skips rows that cannot pass the first test, and the
    listed rows keep the library order: all five outputs are bit-identical with and
    without it."""
    import torch
    from pyebsdindex_amd import tripletvote as tv
    ind = tv.addphase(phase, angTol=tol)
    bn, _ = cubic_bands(ind, 400, 9, seed=91, noise_deg=0.4)
    rng = np.random.default_rng(3)
    inten = rng.uniform(0.1, 1.0, (400, 9)).astype(np.float32)
    inten[rng.random((400, 9)) < 0.05] = 0.0
    dev = torch.device("cuda", 0)
    bnd, intd = torch.as_tensor(bn, device=dev), torch.as_tensor(inten, device=dev)
    a = ind.tripvote(bnd, intd, use_index=True)
    b = ind.tripvote(bnd, intd, use_index=False)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    # an index built for a smaller tolerance would drop matching rows: refused
    st, rw, bw, built = tv.row_index(ind.angtriplets["angles"], tol - 0.25)
    small = (torch.as_tensor(st, device=dev), torch.as_tensor(rw, device=dev), bw, built)
    with pytest.raises(ValueError, match="cannot serve"):
        tv.tripvote_device(bnd, intd, ind.lut, tol, ind.angtriplets["angles"], ind.angtriplets["familyid"],
                           ind.npolefamilies, index=small)


def test_no_patterns():
    """Zero patterns: empty outputs of the right shapes, no launch."""
    import torch
    from pyebsdindex_amd import tripletvote as tv
    ind = tv.addphase("FCC")
    out = ind.tripvote(torch.zeros((0, 9, 3), dtype=torch.float32, device="cuda"))
    assert [tuple(t.shape) for t in out] == [(0, 4, 9), (0, 9), (0, 9), (0, 9), (0, 4, 9)]


def test_vote_on_detected_bands():
    """The chain on the device: synthetic FCC patterns (120x120 uint16, pattern centre
    (0.5, 0.5, 0.6)) -> find_bands -> radon2pole -> vote, against the oracle's vote on the
    same band normals and intensities (avemax of the valid bands, as the indexer passes)."""
    import torch
    from helpers import make_pair
    from pyebsdindex_amd import tripletvote as tv
    from pyebsdindex_amd.differentiable import radon2pole
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    pc = (0.5, 0.5, 0.6)
    pats = kikuchi_patterns(64, (120, 120), "uint16", seed=5, family="FCC", pc=pc).numpy()
    bd, _ = make_pair((120, 120))
    got = bd.find_bands(pats)
    dev = torch.device("cuda", 0)
    theta = torch.as_tensor(np.ascontiguousarray(got["theta"]), device=dev)
    rho = torch.as_tensor(np.ascontiguousarray(got["rho"]), device=dev)
    n = radon2pole(theta, rho, pc, 120, 120)
    inten = (got["avemax"] * got["valid"]).astype(np.float32)
    ind = tv.addphase("FCC")
    (accW, fam, rank, cm, acc), _ = _compare(ind, n.cpu().numpy(), inten)
    valid = got["valid"] > 0
    # a detected band of an FCC pattern is voted for: its best family holds votes from
    # several triplets
    best = acc.max(axis=1)
    assert np.median(best[valid]) >= 5
