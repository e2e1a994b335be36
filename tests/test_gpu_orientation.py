"""Downstream orientations through an independent fit: the device's bands against the FP64
oracle's bands, each band assigned to the generator's known plane normal and the orientation
solved as a weighted Wahba problem by SVD (oracle/orientation.py), must agree within 0.05 deg.
This isolates how band-position differences move an orientation; the reference's own
consumer (vote, band assignment, QUEST) is the chain of tests/test_gpu_bandindex.py."""
import pytest

from helpers import make_pair, oracle64, orientation_agreement

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,dtype,family,n", [((120, 120), "uint16", "FCC", 48),
                                                  ((60, 60), "uint8", "FCC", 48),
                                                  ((160, 160), "uint8", "BCC", 16)])
def test_orientations_match_oracle(shape, dtype, family, n):
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    print(f"orientations {shape} {dtype} {family} n={n}", flush=True)
    pc = (0.5, 0.5, 0.6)
    pats, meta = kikuchi_patterns(n, shape, dtype, seed=61, family=family, pc=pc, return_meta=True)
    pats = pats.numpy()
    bd, ref = make_pair(shape)
    got = bd.find_bands(pats)
    res, _, _, _ = oracle64(ref, pats)
    fitted, worst, med_true = orientation_agreement(got, res, meta, pc, shape)
    print(f"  {fitted}/{n} patterns fitted, device vs oracle worst {worst:.2e} deg, "
          f"device vs true median {med_true:.3f} deg", flush=True)
    assert fitted >= 0.75 * n
    assert worst < 0.05
    # the generator's own orientations: a flipped theta/rho convention would miss by degrees
    assert med_true < 0.5
