"""The NLPAR options of the oracle (oracle/nlpar_ref.py calcnlpar_options) and the host parts
of the device path, pinned on the CPU: stem_scale's round trip, rescale's range, numpy's
mean over patterns as the sequential float32 sum the device restates, and the background
fit (the reference's NLPAR.backsub, nlpar_cpu.py:679-735) on an exact surface."""
import numpy as np
import pytest

from oracle import nlpar_ref as nr


def _scan(nrows, ncols, shape, dtype, seed):
    """Every point its own random pattern: neighbours are far apart in the NLPAR distance,
    so a small lambda gives them no weight."""
    rng = np.random.default_rng(seed)
    top = np.iinfo(dtype).max
    base = rng.uniform(0.2, 0.6, size=(nrows * ncols,) + shape) * top
    return rng.poisson(base).clip(0, top).astype(dtype)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_stem_scale_round_trip(dtype):
    """A scan of one repeated pattern averages to that pattern (weights that sum to one over
    equal values); stem_scale's sqrt and square around the average then return the values to
    within one count (the float32 square of a rounded root truncates to v - 1 at most)."""
    nrows, ncols, shape = 7, 7, (20, 20)
    one = _scan(1, 1, shape, dtype, 3)
    pats = np.repeat(one, nrows * ncols, axis=0)
    idx = np.flatnonzero(nr.makeautomask(*shape).ravel())
    for stem in (False, True):
        out = nr.calcnlpar_options(pats, nrows, ncols, idx, lam=0.7, sr=2, stem_scale=stem, out_dtype=dtype)
        d = pats.astype(int) - out.reshape(pats.shape).astype(int)
        assert d.min() >= 0 and d.max() <= 1, stem
        assert (d == 0).mean() > 0.5, stem


def test_rescale_maps_each_pattern_to_the_range():
    nrows, ncols, shape = 6, 6, (16, 16)
    pats = _scan(nrows, ncols, shape, np.uint8, 4)
    idx = np.flatnonzero(nr.makeautomask(*shape).ravel())
    out = nr.calcnlpar_options(pats, nrows, ncols, idx, lam=0.7, sr=2, rescale_top=65535, out_dtype=np.uint16)
    out = out.reshape(nrows * ncols, -1).astype(int)
    assert np.all(out.min(axis=1) == 0)
    assert np.all(out.max(axis=1) >= 65534)


def test_numpy_mean_over_patterns_is_sequential_float32():
    """np.mean(data, axis=0) of float32 (n, H, W) adds the patterns one after another in
    float32 and divides in float64, rounding to float32 -- the order the device's column
    sums follow (ebsd_nlpar_colsum)."""
    rng = np.random.default_rng(5)
    for n in (1, 7, 300):
        x = (rng.standard_normal((n, 9, 11)) * 1000).astype(np.float32)
        acc = x[0].copy()
        for i in range(1, n):
            acc = acc + x[i]
        want = (acc.astype(np.float64) / n).astype(np.float32)
        assert np.array_equal(np.mean(x, axis=0), want), n


def _surface(ny, nx, a=120.0, x0=23.5, y0=17.25, sx=9.0, sy=7.5, c=40.0, d=0.3, e=-0.2):
    yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
    return a * np.exp(-((xx - x0) ** 2) / (2 * sx ** 2) - ((yy - y0) ** 2) / (2 * sy ** 2)) + c + d * xx + e * yy


def test_background_fit_recovers_exact_surface():
    """A mean pattern that is exactly a Gaussian on a plane is recovered by the fit (the
    oracle's and the device path's host fit alike), shifted to zero mean."""
    from pyebsdindex_amd.nlpar import fit_background, makeautomask
    ny, nx = 40, 48
    s = _surface(ny, nx)
    mask = makeautomask(ny, nx)
    want = s - s.mean()
    for fit in (nr.backsub_fit(s, mask), fit_background(s, mask)):
        assert fit.dtype == np.float64
        np.testing.assert_allclose(fit, want, atol=1e-6)
        assert abs(fit.mean()) < 1e-9
    # float32 mean patterns (as the scan's mean is) fit the same surface to float32 accuracy
    np.testing.assert_allclose(fit_background(s.astype(np.float32), mask), want, atol=2e-3)


def test_backsub_removes_a_shared_background():
    """Patterns = one surface + small independent noise: after backsub the written
    (averaged) patterns are flat around the surface's mean level."""
    nrows, ncols, shape = 6, 6, (24, 24)
    rng = np.random.default_rng(6)
    s = _surface(*shape, a=60.0, x0=11.0, y0=12.5, sx=6.0, sy=5.0, c=100.0, d=0.2, e=0.1)
    pats = (s + rng.normal(0, 1.0, (nrows * ncols,) + shape)).clip(0, 255).astype(np.uint8)
    mask = nr.makeautomask(*shape)
    idx = np.flatnonzero(mask.ravel())
    out = nr.calcnlpar_options(pats, nrows, ncols, idx, lam=0.7, sr=2, backsub=True, mask=mask)
    flat = out.reshape(nrows * ncols, *shape).astype(float)
    inside = mask > 0
    # the Gaussian's 60-count peak is gone: the spread inside the mask is the noise's
    assert flat[:, inside].std() < 3.0 and np.ptp(s[inside]) > 50
