"""Pins of the orientation checker (oracle/orientation.py): the two Wahba solutions against
known rotations and against each other, and the band-to-plane assignment."""
import numpy as np

from oracle import orientation as ori
from pyebsdindex_amd.synthetic import plane_normals, quat_to_matrix, random_quaternions


def _case(seed, m=9, noise=0.0):
    rng = np.random.default_rng(seed)
    R = quat_to_matrix(random_quaternions(1, rng))[0]
    c = rng.standard_normal((m, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    n = c @ R.T + noise * rng.standard_normal((m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    w = rng.uniform(0.5, 2.0, m)
    return R, n, c, w


def test_exact_vectors_recover_rotation():
    """Noise-free pairs n = R c fix R: both solutions return it to rounding, and a transposed
    or un-rotated answer would fail (the rotations are far from symmetric)."""
    for seed in range(20):
        R, n, c, w = _case(seed)
        for solve in (ori.wahba_svd, ori.wahba_davenport):
            Rs = solve(n, c, w)
            assert np.abs(Rs - R).max() < 1e-12, (seed, solve.__name__)
            assert np.allclose(Rs @ Rs.T, np.eye(3), atol=1e-12)
            assert np.linalg.det(Rs) > 0
        assert np.abs(R - R.T).max() > 1e-3


def test_svd_and_davenport_agree_with_noise():
    """Two textbook solutions of the same weighted problem agree on noisy vectors, and the
    answer minimises the loss (no nearby rotation does better)."""
    for seed in range(20):
        R, n, c, w = _case(100 + seed, noise=0.01)
        Ra = ori.wahba_svd(n, c, w)
        Rb = ori.wahba_davenport(n, c, w)
        assert ori.misorientation_deg(Ra, Rb) < 1e-9
        loss = lambda Q: float(np.sum(w * np.sum((n - c @ Q.T) ** 2, axis=1)))
        rng = np.random.default_rng(seed)
        for _ in range(10):
            ax = rng.standard_normal(3)
            ax /= np.linalg.norm(ax)
            K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
            dR = np.eye(3) + np.sin(1e-3) * K + (1 - np.cos(1e-3)) * K @ K
            assert loss(dR @ Ra) > loss(Ra)


def test_misorientation_of_known_rotation():
    th = np.radians(0.05)
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    R = quat_to_matrix(random_quaternions(1, np.random.default_rng(4)))[0]
    assert abs(ori.misorientation_deg(R, R @ Rz) - 0.05) < 1e-9


def test_assignment_recovers_planes_and_signs():
    """Detected normals of random sign, 0.5 deg off their planes, go to the right plane
    with the sign that matches; normals far from every plane and invalid slots are dropped."""
    rng = np.random.default_rng(7)
    c = plane_normals("FCC")
    R = quat_to_matrix(random_quaternions(1, rng))[0]
    g = c @ R.T
    pick = rng.choice(len(c), 6, replace=False)
    sg = rng.choice([-1.0, 1.0], 6)
    nd = sg[:, None] * g[pick]
    tilt = rng.standard_normal((6, 3))
    tilt -= np.sum(tilt * nd, 1, keepdims=True) * nd
    tilt /= np.linalg.norm(tilt, axis=1, keepdims=True)
    nd = np.cos(np.radians(0.5)) * nd + np.sin(np.radians(0.5)) * tilt
    valid = np.array([1, 1, 1, 1, 1, 0])
    n, cm, idx = ori.assign_bands(nd, g, c, valid, max_deg=3.0)
    assert list(idx) == [0, 1, 2, 3, 4]
    assert np.allclose(cm, sg[:5, None] * c[pick[:5]])
    Rf = ori.wahba_svd(n, cm)
    assert ori.misorientation_deg(Rf, R) < 1.0


def test_reference_restatement_orientations_match_fp64():
    """The float32 reference restatement (oracle/ref_cpu.py: numba loops in C, scipy FFT
    convolution) against the FP64 oracle: orientations within 0.05 deg, and both near the
    generator's own orientations."""
    from helpers import make_pair, oracle64, orientation_agreement
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    shape, pc = (60, 60), (0.5, 0.5, 0.6)
    pats, meta = kikuchi_patterns(24, shape, "uint8", seed=61, pc=pc, return_meta=True)
    pats = pats.numpy()
    _, ref = make_pair(shape)
    got = ref.find_bands(pats)
    res, _, _, _ = oracle64(ref, pats)
    fitted, worst, med_true = orientation_agreement(got, res, meta, pc, shape)
    assert fitted >= 18
    assert worst < 0.05
    assert med_true < 0.5
