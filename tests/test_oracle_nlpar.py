"""Pins of the NLPAR oracle (oracle/nlpar_ref.py): the C restatement of the reference's
Numba kernels (nlpar_cpu.py:753-935) and its float64 form, against closed forms."""
import numpy as np
import pytest

from oracle import nlpar_ref as nr


def _scan(nrows, ncols, H, W, seed, scale=50.0, base=100.0):
    rng = np.random.default_rng(seed)
    return (base + scale * rng.standard_normal((nrows * ncols, H * W))).astype(np.float32)


def test_automask_is_the_centred_disc():
    """For even sizes the rolled mask is the disc of radius 0.49 min(h, w) about (h/2, w/2)."""
    for h, w in ((60, 60), (80, 100), (120, 120)):
        m = nr.makeautomask(h, w)
        y, x = np.mgrid[0:h, 0:w]
        assert np.array_equal(m.astype(bool), np.hypot(y - h / 2, x - w / 2) < 0.49 * min(h, w))


def test_sigma_of_constant_offsets():
    """Neighbours that differ from the centre by a constant c over the mask give
    d^2 / 2n = c^2 / 2, so sigma is the smallest |c| / sqrt(2); an identical neighbour
    (d^2 < 1e-3, a pattern recorded twice) does not count."""
    H = W = 6
    rng = np.random.default_rng(0)
    A = (100 + 10 * rng.standard_normal(H * W)).astype(np.float32)
    offs = [3.0, -2.0, 5.0, 0.0, 7.0, -4.0, 6.0, 2.5]
    data = np.empty((9, H * W), np.float32)
    k = 0
    for p in range(9):
        if p == 4:
            data[p] = A
        else:
            data[p] = A + np.float32(offs[k])
            k += 1
    idx = np.arange(H * W)
    for fn in (lambda: nr.sigma(data, 3, 3, idx, saturation_protect=False)[0],
               lambda: nr.sigma64(data, 3, 3, idx, saturation_protect=False)):
        s = fn()
        assert abs(s[1, 1] - 2.0 / np.sqrt(2.0)) < 1e-5


def test_self_entry_of_normalised_distances():
    """The centre entry has d^2 = 0 and count P, so its normalised distance is
    (0 - P 2 s^2) / (2 s^2 sqrt(2P)) = -sqrt(P/2) whatever sigma is."""
    data = _scan(4, 5, 7, 7, seed=1)
    s, dout, nout = nr.sigma(data, 4, 5, np.arange(49), nn=1)
    centre = dout[:, :, 4][1:-1, 1:-1]             # interior points: the centre is entry 4
    assert np.allclose(centre, -np.sqrt(49 / 2.0), rtol=1e-6)
    assert np.all(nout[1:-1, 1:-1, 4] == 49)
    # a corner point's clipped window has 4 entries; the rest stay 0
    assert np.all(nout[0, 0, 4:] == 0) and np.all(nout[0, 0, :4] > 0)


def test_saturated_pixels_drop_out():
    """Pixels at or above 0.9961 x the scan maximum are left out of the pair counts."""
    data = _scan(3, 3, 8, 8, seed=2)
    top = data.max() + 100.0                       # well above every other pixel
    data[0, :5] = top                              # 5 saturated pixels in pattern 0
    data[1, 3:9] = top                             # 6 in pattern 1 (2 shared: 3, 4)
    idx = np.arange(64)
    _, _, nout = nr.sigma(data, 3, 3, idx, nn=1, saturation_protect=True)
    # point (0,0): entries (0,0) self, (0,1) = pattern 1, (1,0) = pattern 3, (1,1) = pattern 4
    assert nout[0, 0, 1] == 64 - 9                 # union of the two saturated sets
    assert nout[0, 0, 2] == 64 - 5
    _, _, nout2 = nr.sigma(data, 3, 3, idx, nn=1, saturation_protect=False)
    assert nout2[0, 0, 1] == 64


@pytest.mark.parametrize("nrows,ncols,sr", [(7, 9, 3), (5, 6, 2), (3, 3, 1)])
def test_large_lambda_gives_window_mean(nrows, ncols, sr):
    """lambda -> infinity sends every weight to 1: each point becomes the mean of its
    (2sr+1)^2 window, shifted (not clipped) at the scan edges."""
    H = W = 5
    data = _scan(nrows, ncols, H, W, seed=3)
    idx = np.arange(H * W)
    sig = nr.sigma(data, nrows, ncols, idx)[0]
    out = nr.nlpar(data, nrows, ncols, idx, sig, lam=1e6, sr=sr)
    out64 = nr.nlpar64(data, nrows, ncols, idx, sig, lam=1e6, sr=sr)
    cube = data.reshape(nrows, ncols, -1).astype(np.float64)
    ws = 2 * sr + 1
    for j in range(nrows):
        for i in range(ncols):
            y0 = int(np.clip(j - sr, 0, nrows - ws))
            x0 = int(np.clip(i - sr, 0, ncols - ws))
            want = cube[y0:y0 + ws, x0:x0 + ws].reshape(-1, H * W).mean(axis=0)
            assert np.allclose(out[j * ncols + i], want, rtol=2e-6), (j, i)
            assert np.allclose(out64[j * ncols + i], want, rtol=1e-12), (j, i)


def test_small_lambda_gives_the_pattern():
    """lambda -> 0 with distinct patterns and small sigmas leaves only the centre weight."""
    data = _scan(5, 5, 6, 6, seed=4)
    idx = np.arange(36)
    sig = np.full((5, 5), 0.05, np.float32)
    out = nr.nlpar(data, 5, 5, idx, sig, lam=1e-3, sr=2)
    assert np.array_equal(out, data)


def test_c_restatement_against_float64():
    """The float32 restatement agrees with the float64 definitions to float32 rounding,
    including a mask, saturation, a search radius of 2 and dthresh."""
    nrows, ncols, H, W = 6, 7, 10, 10
    rng = np.random.default_rng(5)
    base = rng.uniform(50, 200, H * W)
    data = (base[None] + 20 * rng.standard_normal((nrows * ncols, H * W))).clip(0, 255).astype(np.float32)
    idx = np.nonzero(nr.makeautomask(H, W).ravel())[0]
    s32 = nr.sigma(data, nrows, ncols, idx)[0]
    s64 = nr.sigma64(data, nrows, ncols, idx)
    assert np.allclose(s32, s64, rtol=1e-5)
    o32 = nr.nlpar(data, nrows, ncols, idx, s32, lam=0.9, sr=2, dthresh=0.1)
    o64 = nr.nlpar64(data, nrows, ncols, idx, s32, lam=0.9, sr=2, dthresh=0.1)
    assert np.allclose(o32, o64, rtol=2e-4, atol=1e-3)


def test_calclim_restricts_the_points():
    data = _scan(8, 8, 4, 4, seed=6)
    idx = np.arange(16)
    sig = nr.sigma(data, 8, 8, idx)[0]
    full = nr.nlpar(data, 8, 8, idx, sig, sr=2)
    part = nr.nlpar(data, 8, 8, idx, sig, sr=2, calclim=(2, 1, 3, 4)).reshape(8, 8, -1)
    f = full.reshape(8, 8, -1)
    assert np.array_equal(part[1:5, 2:5], f[1:5, 2:5])
    part[1:5, 2:5] = 0
    assert not part.any()
