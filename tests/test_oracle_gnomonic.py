"""The oracle's radon2pole restatement against what the geometry fixes."""
import numpy as np

from oracle import gnomonic


def test_radon2pole_normals_are_perpendicular_to_band_and_pc_ray():
    rng = np.random.default_rng(3)
    N, nB, H, W = 5, 7, 120, 160
    theta = rng.uniform(0.05, np.pi, (N, nB)).astype(np.float32)
    rho = rng.uniform(-50, 50, (N, nB)).astype(np.float32)
    for vendor in ("EDAX", "OXFORD", "BRUKER"):
        pc = rng.uniform([0.4, 0.4, 0.5], [0.6, 0.6, 0.8], (N, 3))
        n = gnomonic.radon2pole(theta, rho, pc, (H, W), vendor).astype(np.float64)
        t = gnomonic.vendor_t(pc, (H, W), vendor).astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(n, axis=-1), 1.0, atol=1e-6)
        th, rh = theta.astype(np.float64), rho.astype(np.float64)
        d = np.stack([-np.sin(th), np.cos(th), np.zeros_like(th)], -1)      # band direction
        for s in (-30.0, 0.0, 45.0):                                          # points on the band line
            P = np.stack([rh * np.cos(th) + W / 2 + s * d[..., 0], rh * np.sin(th) + H / 2 + s * d[..., 1],
                          np.zeros_like(th)], -1)
            v = P - t[:, None, :]
            v /= np.linalg.norm(v, axis=-1, keepdims=True)
            assert np.abs((n * v).sum(-1)).max() < 2e-6
        assert np.abs((n * d).sum(-1)).max() < 2e-6


def test_vendor_translation_matches_reference_conventions():
    # EDAX: t = (x* W, y* H, -z* min(H, W)); Bruker: y measured from the top
    t = gnomonic.vendor_t(np.array([0.5, 0.25, 0.75]), (100, 200), "EDAX")
    np.testing.assert_allclose(t, [100.0, 25.0, -75.0])
    t = gnomonic.vendor_t(np.array([0.5, 0.25, 0.75]), (100, 200), "BRUKER")
    np.testing.assert_allclose(t, [100.0, 75.0, -75.0])
    t = gnomonic.vendor_t(np.array([0.5, 0.25, 0.75]), (100, 200), "OXFORD")
    np.testing.assert_allclose(t, [100.0, 50.0, -150.0])
