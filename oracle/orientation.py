"""Orientation from detected bands (test infrastructure: the checker for the downstream
orientation bar, never the product path).

The reference turns band normals into orientations with triplet voting followed by a
QUEST refinement (pyebsdindex/tripletvote.py:492-633; the weighted fit at
``_refine_orientation_quest``).  Triplet voting needs numba and is not run here.  The
checks in ``tests/`` therefore assign each detected band to the generator's known
plane normal (the one closest to it) and solve the same weighted least-squares problem the
QUEST step solves -- Wahba's problem, min_R sum_i w_i |n_i - R c_i|^2 -- in two
independent textbook forms:

* ``wahba_svd``: Markley's SVD solution (J. Astronaut. Sci. 36, 1988): B = sum w n c^T,
  R = U diag(1, 1, det U det V) V^T.
* ``wahba_davenport``: Davenport's q-method: the rotation quaternion is the eigenvector of
  the largest eigenvalue of the 4x4 K matrix built from B.

Both run in float64.  Orientations from the device bands and from the FP64 oracle's bands
then differ only by what the band positions differ.
"""
import numpy as np


def wahba_svd(n, c, w=None):
    """Rotation R (3x3) minimising sum_i w_i |n_i - R c_i|^2 (n, c: (M, 3))."""
    n = np.asarray(n, np.float64)
    c = np.asarray(c, np.float64)
    w = np.ones(n.shape[0]) if w is None else np.asarray(w, np.float64)
    B = np.einsum("i,ij,ik->jk", w, n, c)
    U, _, Vt = np.linalg.svd(B)
    d = np.linalg.det(U) * np.linalg.det(Vt)
    return U @ np.diag([1.0, 1.0, d]) @ Vt


def wahba_davenport(n, c, w=None):
    """Same problem through Davenport's K matrix; returns R (3x3)."""
    n = np.asarray(n, np.float64)
    c = np.asarray(c, np.float64)
    w = np.ones(n.shape[0]) if w is None else np.asarray(w, np.float64)
    B = np.einsum("i,ij,ik->jk", w, n, c)
    S = B + B.T
    sig = np.trace(B)
    z = np.array([B[1, 2] - B[2, 1], B[2, 0] - B[0, 2], B[0, 1] - B[1, 0]])
    K = np.zeros((4, 4))
    K[0, 0] = sig
    K[0, 1:] = z
    K[1:, 0] = z
    K[1:, 1:] = S - sig * np.eye(3)
    vals, vecs = np.linalg.eigh(K)
    q = vecs[:, np.argmax(vals)]              # (w, x, y, z) of R^T for this sign of z
    w0, x, y, zz = q[0], -q[1], -q[2], -q[3]
    return np.array([
        [1 - 2 * (y * y + zz * zz), 2 * (x * y - w0 * zz), 2 * (x * zz + w0 * y)],
        [2 * (x * y + w0 * zz), 1 - 2 * (x * x + zz * zz), 2 * (y * zz - w0 * x)],
        [2 * (x * zz - w0 * y), 2 * (y * zz + w0 * x), 1 - 2 * (x * x + y * y)],
    ])


def misorientation_deg(Ra, Rb):
    """Rotation angle (degrees) of Ra^T Rb (no crystal symmetry: both fits use the same
    assignment, so they describe the same variant)."""
    M = np.asarray(Ra).T @ np.asarray(Rb)
    cosang = (np.trace(M) - 1.0) / 2.0
    # sin from the skew part keeps small angles accurate (arccos near 1 loses half the digits)
    sinang = 0.5 * np.linalg.norm([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    return float(np.degrees(np.arctan2(sinang, cosang)))


def assign_bands(normals_det, g_det, crystal_normals, valid, max_deg=3.0):
    """Assign each valid detected normal (nB, 3) to the closest known plane normal
    ``g_det`` (K, 3, detector frame) up to sign, if within ``max_deg``.  Returns
    (detected normals, sign-matched crystal normals, band indices)."""
    n = np.asarray(normals_det, np.float64)
    n = n / np.linalg.norm(n, axis=1, keepdims=True)
    dots = n @ np.asarray(g_det, np.float64).T          # (nB, K)
    best = np.argmax(np.abs(dots), axis=1)
    cosb = np.abs(dots[np.arange(n.shape[0]), best])
    ok = (np.asarray(valid) > 0) & (cosb >= np.cos(np.radians(max_deg)))
    idx = np.nonzero(ok)[0]
    sgn = np.sign(dots[idx, best[idx]])
    return n[idx], sgn[:, None] * np.asarray(crystal_normals, np.float64)[best[idx]], idx
