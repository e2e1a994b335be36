"""Gnomonic band geometry -- TEST INFRASTRUCTURE ONLY.

``radon2pole`` restates ``Radon.radon2pole`` (pyebsdindex/radon_fast.py:314-397) for the
three-parameter vendors (EDAX, OXFORD, BRUKER/KIKUCHIPY) with the reference's float32
arrays.  It is pinned in tests/test_oracle_gnomonic.py by what the geometry fixes: the
normal is a unit vector perpendicular to the band direction and to the vector from the
pattern centre to every point of the band line on the detector.
"""
import numpy as np


def vendor_t(PC, dim, vendor):
    """Detector-frame translation t of the pattern centre (radon_fast.py:343-370)."""
    dimf = np.array(dim, dtype=np.float32)
    t = np.asarray(PC, dtype=np.float32).copy()
    ven = str(vendor).upper()
    if ven == 'EDAX':
        t *= np.array([dimf[1], dimf[0], -np.min(dimf[0:2])])
    elif ven == 'OXFORD':
        t *= np.array([dimf[1], dimf[1], -dimf[1]])
    elif ven in ('KIKUCHIPY', 'BRUKER'):
        t *= np.array([dimf[1], dimf[0], -dimf[0]])
        t[..., 1] = dimf[0] - t[..., 1]
    else:
        raise ValueError(vendor)
    return t


def radon2pole(theta, rho, PC, dim, vendor='EDAX'):
    """(N, nB) band theta, rho -> (N, nB, 3) unit normals (radon_fast.py:335-397)."""
    theta = np.asarray(theta, dtype=np.float32)
    rho = np.asarray(rho, dtype=np.float32)
    nPats, nBands = theta.shape
    stheta = np.sin(theta)
    ctheta = np.cos(theta)
    pct = np.asarray(PC, dtype=np.float32)
    if pct.ndim < 2:
        pct = np.tile(pct, nPats).reshape(nPats, 3)
    t = vendor_t(pct, dim, vendor)
    dimf = np.array(dim, dtype=np.float32)
    t = np.tile(t.reshape(nPats, 1, 3), (1, nBands, 1))
    r = np.zeros((nPats, nBands, 3), dtype=np.float32)
    r[:, :, 0] = -1 * stheta
    r[:, :, 1] = ctheta
    p = np.zeros((nPats, nBands, 3), dtype=np.float32)
    p[:, :, 0] = rho * ctheta
    p[:, :, 1] = rho * stheta
    p[:, :, 0] += dimf[1] * 0.5
    p[:, :, 1] += dimf[0] * 0.5
    n2 = p - t
    n = np.cross(r.reshape(nPats * nBands, 3), n2.reshape(nPats * nBands, 3))
    norm = np.linalg.norm(n, axis=1).clip(1e-12)
    n /= norm.reshape(nPats * nBands, 1)
    return n.reshape(nPats, nBands, 3)
