"""Adjoint kernels against FP64 finite differences (central, step 1e-6) and the reference
geometry; pattern-centre refinement against the hidden centre the patterns were made at."""
import numpy as np
import pytest

from helpers import make_pair
from oracle import fp64, gnomonic

pytestmark = pytest.mark.gpu

H_STEP = 1e-6


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def test_backprojection_is_the_radon_adjoint():
    """<v, A^T y> against the FP64 central difference of <radon64(x + h v), y>."""
    from pyebsdindex_amd.differentiable import backproject
    rng = np.random.default_rng(1)
    shape = (60, 80)
    bd, ref = make_pair(shape)
    plan = bd.device_plan()
    x = rng.uniform(0, 200, (2,) + shape)
    v = rng.standard_normal((2,) + shape)
    y = rng.standard_normal((2, bd.nRho, bd.nTheta))
    dX = backproject(plan, _dev(y.astype(np.float32))).cpu().numpy().astype(np.float64)
    plan64 = ref.radonPlan.indexPlan
    f = lambda z: float((fp64.radon64(z, plan64) * y).sum())
    fd = (f(x + H_STEP * v) - f(x - H_STEP * v)) / (2 * H_STEP)
    an = float((dX * v).sum())
    assert abs(an - fd) <= 1e-4 * abs(fd), (an, fd)
    # pixel by pixel on a few pixels (exactly linear: the dot-product identity)
    y32 = y.astype(np.float32).astype(np.float64)
    zero = fp64.radon64(np.zeros_like(x), plan64)  # -1 on lines without pixels: not linear there
    for q, i, j in [(0, 0, 0), (0, 30, 40), (1, 59, 79), (1, 10, 70)]:
        e = np.zeros_like(x)
        e[q, i, j] = 1.0
        want = float(((fp64.radon64(e, plan64) - zero) * y32).sum())
        assert abs(dX[q, i, j] - want) <= 1e-5 * max(1.0, abs(want))


def test_conv_transpose_is_the_padding_convolution_adjoint():
    from pyebsdindex_amd.differentiable import conv_transpose
    rng = np.random.default_rng(2)
    bd, ref = make_pair((60, 60))
    plan = bd.device_plan()
    p0, p1 = (int(v) for v in ref.padding)
    R = rng.standard_normal((2, bd.nRho, bd.nTheta))
    V = rng.standard_normal(R.shape)
    Y = rng.standard_normal((2, bd.nRho + 2 * p0, bd.nTheta + 2 * p1))
    dR = conv_transpose(plan, _dev(Y.astype(np.float32))).cpu().numpy().astype(np.float64)
    K = ref.kernel[0]
    f = lambda z: float((fp64.conv64(fp64.pad64(z, p0, p1), K) * Y).sum())
    fd = (f(R + H_STEP * V) - f(R - H_STEP * V)) / (2 * H_STEP)
    an = float((dR * V).sum())
    assert abs(an - fd) <= 1e-4 * abs(fd), (an, fd)


def test_band_position_gradient_vs_fp64_finite_differences():
    """d(sum w_rho rho + w_theta theta)/d(pattern) along a random direction: the device
    backward (band-fit Jacobian, transposed convolution, backprojection) against the FP64
    oracle's central difference.  Bands whose integer bins differ from the oracle's (exact
    ties) are left out of the loss."""
    import torch
    from pyebsdindex_amd.differentiable import BandPeaks
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    shape = (120, 120)
    bd, ref = make_pair(shape)
    pats = kikuchi_patterns(4, shape, "float32", seed=61).numpy().astype(np.float64)
    x = _dev(pats.astype(np.float32)).requires_grad_(True)
    rho, theta, valid = BandPeaks.apply(x, bd)
    res, *_ = fp64.find_bands64(pats.astype(np.float32).astype(np.float64), ref.radonPlan.indexPlan,
                                ref.kernel[0], ref.rdnmask, ref.padding, ref.peakPad, ref.nBands,
                                ref.radonPlan.theta, ref.radonPlan.rho)
    rng = np.random.default_rng(5)
    w_r = rng.standard_normal(rho.shape)
    w_t = rng.standard_normal(rho.shape) * 30.0
    # bands used: valid in both, with the oracle's integer bins
    dev_loc = bd.find_bands_device(x.detach())['maxloc'].cpu().numpy()
    same = (valid.cpu().numpy() == 1) & (res['valid'] == 1) & np.all(dev_loc == res['maxloc'], -1)
    keep = same.astype(np.float64)
    L = (torch.as_tensor(w_r * keep, device=x.device) * rho.double()).sum() + \
        (torch.as_tensor(w_t * keep, device=x.device) * theta.double()).sum()
    L.backward()
    g = x.grad.cpu().numpy().astype(np.float64)
    v = rng.standard_normal(pats.shape) * pats.std()   # a direction on the intensity scale
    base = pats.astype(np.float32).astype(np.float64)

    def f(z):
        r, *_ = fp64.find_bands64(z, ref.radonPlan.indexPlan, ref.kernel[0], ref.rdnmask, ref.padding,
                                  ref.peakPad, ref.nBands, ref.radonPlan.theta, ref.radonPlan.rho)
        assert np.array_equal(r['maxloc'], res['maxloc'])
        return float((w_r * keep * r['rho']).sum() + (w_t * keep * r['theta']).sum())

    fd = (f(base + H_STEP * v) - f(base - H_STEP * v)) / (2 * H_STEP)
    an = float((g * v).sum())
    assert keep.sum() >= 20
    assert abs(an - fd) <= 1e-4 * abs(fd), (an, fd)


def test_gnomonic_bands_invert_radon2pole_and_pc_jacobian_vs_fd():
    """Predicted bands fed to the reference's radon2pole give back the normals; the
    d rho / d PC Jacobian matches FP64 central differences of the prediction."""
    import torch
    from pyebsdindex_amd.differentiable import gnomonic_bands
    rng = np.random.default_rng(7)
    N, M, H, W = 3, 11, 120, 140
    n = rng.standard_normal((N, M, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    for vendor in ("EDAX", "OXFORD", "BRUKER"):
        pc = np.array([0.47, 0.53, 0.62])
        rho, th, J = (t.cpu().numpy() for t in gnomonic_bands(_dev(n), pc, H, W, vendor))
        back = gnomonic.radon2pole(th, rho, pc, (H, W), vendor).astype(np.float64)
        cos = np.abs((back * n).sum(-1))
        assert cos.min() > 1 - 1e-6
        for k in range(3):
            e = np.zeros(3)
            e[k] = H_STEP
            rp = gnomonic_bands(_dev(n), pc + e, H, W, vendor)[0].cpu().numpy()
            rm = gnomonic_bands(_dev(n), pc - e, H, W, vendor)[0].cpu().numpy()
            fd = (rp - rm) / (2 * H_STEP)
            np.testing.assert_allclose(J[..., k], fd, rtol=1e-4, atol=1e-4 * np.abs(fd).max())


def test_radon2pole_device_matches_reference_restatement():
    from pyebsdindex_amd.differentiable import radon2pole
    rng = np.random.default_rng(9)
    N, nB, H, W = 4, 9, 120, 120
    theta = rng.uniform(0.02, np.pi, (N, nB)).astype(np.float32)
    rho = rng.uniform(-40, 40, (N, nB)).astype(np.float32)
    pc = rng.uniform([0.45, 0.45, 0.55], [0.55, 0.55, 0.7], (N, 3))
    for vendor in ("EDAX", "OXFORD", "BRUKER"):
        got, J = radon2pole(_dev(theta), _dev(rho), pc, H, W, vendor, jacobian=True)
        want = gnomonic.radon2pole(theta, rho, pc, (H, W), vendor)
        np.testing.assert_allclose(got.cpu().numpy(), want, atol=2e-6)
        J = J.cpu().numpy()
        for k in range(3):
            e = np.zeros(3)
            e[k] = H_STEP
            fp = gnomonic.radon2pole(theta, rho, pc + e, (H, W), vendor).astype(np.float64)
            # float64 finite difference of the same geometry (the float32 restatement is too coarse)
            fd = (_r2p64(theta, rho, pc + e, H, W, vendor) - _r2p64(theta, rho, pc - e, H, W, vendor)) / (2 * H_STEP)
            np.testing.assert_allclose(J[..., k], fd, rtol=1e-4, atol=1e-6)
            del fp


def _r2p64(theta, rho, pc, H, W, vendor):
    th = theta.astype(np.float64)
    rh = rho.astype(np.float64)
    t = np.asarray(pc, dtype=np.float64) * {"EDAX": [W, H, -min(H, W)], "OXFORD": [W, W, -W],
                                              "BRUKER": [W, H, -H]}[vendor]
    if vendor == "BRUKER":
        t[..., 1] = H - t[..., 1]
    r = np.stack([-np.sin(th), np.cos(th), np.zeros_like(th)], -1)
    p = np.stack([rh * np.cos(th) + 0.5 * W, rh * np.sin(th) + 0.5 * H, np.zeros_like(th)], -1)
    n = np.cross(r, p - t[:, None, :])
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


def test_pc_loss_gradient_vs_fp64_fd_and_refinement():
    """256-pattern workload in miniature: float32 patterns at a hidden centre, the loss of
    detected vs predicted bands; its PC gradient against FP64 central differences; and
    Gauss-Newton refinement from a centre perturbed by 0.01."""
    import torch
    from pyebsdindex_amd.differentiable import BandPeaks, gnomonic_bands, match_bands, pc_loss, refine_pc
    from pyebsdindex_amd.synthetic import kikuchi_patterns
    shape = (120, 120)
    bd, _ = make_pair(shape)
    true = np.array([0.512, 0.487, 0.634])
    pats, meta = kikuchi_patterns(32, shape, "float32", seed=71, family="FCC", pc=true, return_meta=True)
    x = pats.cuda()
    normals = torch.as_tensor(meta["g_det"], device=x.device)
    rho, theta, valid = BandPeaks.apply(x, bd)
    pc0 = true + np.array([0.01, -0.01, 0.01])
    rh, th, _ = gnomonic_bands(normals, pc0, 120, 120)
    idx, sgn, mask = match_bands(rho, theta, valid, rh, th, bd.rhoMax)
    assert mask.float().mean() > 0.8
    p = torch.tensor(pc0, dtype=torch.float64, device=x.device, requires_grad=True)
    L = pc_loss(rho, theta, normals, p, 120, 120, idx, sgn, mask)
    (g,) = torch.autograd.grad(L, p)
    g = g.cpu().numpy()
    for k in range(3):
        e = np.zeros(3)
        e[k] = H_STEP
        Lp = float(pc_loss(rho, theta, normals, torch.tensor(pc0 + e, device=x.device), 120, 120, idx, sgn, mask))
        Lm = float(pc_loss(rho, theta, normals, torch.tensor(pc0 - e, device=x.device), 120, 120, idx, sgn, mask))
        fd = (Lp - Lm) / (2 * H_STEP)
        assert abs(g[k] - fd) <= 1e-4 * abs(fd), (k, g[k], fd)
    pc, hist = refine_pc(bd, x, normals, pc0)
    err0 = np.abs(pc0 - true).max()
    err = np.abs(pc.cpu().numpy() - true).max()
    assert hist[-1][1] < hist[0][1]
    assert err < 0.2 * err0, (err, err0, hist)
