"""Differentiable band detection and pattern-centre refinement.

The reference computes band positions but no derivatives.  Two ``torch.autograd.Function``
wrappers expose the adjoint kernels of the C ABI:

* ``BandPeaks``: patterns (N, H, W) float32 -> band (rho, theta) (N, nB) as
  ``BandDetect.find_bands`` computes them (pyebsdindex/band_detect.py:411-526).  Its
  backward recomputes the Radon of the saved patterns (the Radon and its convolution are
  not stored, which keeps memory at the input's size for large batches), then applies the
  Jacobian of the 3x3 sub-pixel fit (``ebsd_band_grad``), the transposed convolution with
  the padding adjoint (``ebsd_conv_transpose``) and the Radon backprojection
  (``ebsd_backproject``).
* ``GnomonicBands``: detector-frame plane normals and the pattern centre -> predicted band
  (rho, theta), the inverse of ``Radon.radon2pole`` (pyebsdindex/radon_fast.py:314-397);
  its backward uses d rho / d PC from ``ebsd_gnomonic_bands``.

``refine_pc`` fits the pattern centre to the detected bands of patterns whose orientations
are known, by Gauss-Newton steps on the gradient and Jacobian of those kernels -- a
gradient-based alternative to the simplex/particle-swarm search of pyebsdindex/pcopt.py.
"""
import ctypes

import numpy as np

from . import _hip

VENDORS = {"EDAX": 0, "OXFORD": 1, "BRUKER": 2, "KIKUCHIPY": 2}


def _vendor(v):
    try:
        return VENDORS[str(v).upper()]
    except KeyError:
        raise ValueError(f"vendor {v!r}: one of {sorted(VENDORS)}") from None


def _stream(dev):
    import torch
    return torch.cuda.current_stream(dev).cuda_stream


class BandPeaks:
    """``BandPeaks.apply(patterns, bd)`` -> (rho, theta, valid).  rho/theta are
    differentiable w.r.t. float32 patterns; valid (int8) is not."""

    @staticmethod
    def apply(pats, bd):
        return _BandPeaksFn.apply(pats, bd)


def _band_peaks_fn():
    import torch

    class Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, pats, bd):
            plan = bd.device_plan(pats.device.index)
            p3 = pats.detach().unsqueeze(0) if pats.dim() == 2 else pats.detach()
            x, flags = bd._prepare(plan, p3)
            out = plan.find_bands(x, flags=flags)
            ctx.bd, ctx.flags, ctx.shape, ctx.dtype = bd, flags, pats.shape, pats.dtype
            ctx.save_for_backward(x, out["maxloc"], out["aveloc"], out["valid"])
            ctx.mark_non_differentiable(out["valid"])
            return out["rho"], out["theta"], out["valid"]

        @staticmethod
        def backward(ctx, g_rho, g_theta, g_valid):
            x, maxloc, aveloc, valid = ctx.saved_tensors
            bd = ctx.bd
            plan = bd.device_plan(x.device.index)
            N = x.shape[0]
            if g_rho is None:
                g_rho = torch.zeros_like(aveloc[..., 0])
            if g_theta is None:
                g_theta = torch.zeros_like(aveloc[..., 0])
            g_rho = g_rho.detach().to(torch.float32).contiguous()
            g_theta = g_theta.detach().to(torch.float32).contiguous()
            dX = backward_bands(plan, x, ctx.flags, maxloc, aveloc, valid, g_rho, g_theta)
            return dX.reshape(ctx.shape).to(ctx.dtype), None

    return Fn


def backward_bands(plan, x, flags, maxloc, aveloc, valid, g_rho, g_theta):
    """d(loss)/d(patterns) from d(loss)/d(rho), d(loss)/d(theta) of find_bands' bands:
    Radon recomputed, then band-fit Jacobian, transposed convolution, backprojection."""
    import torch
    lib = plan.lib
    N = x.shape[0]
    nRp = plan.nR + 2 * int(plan.padding[0])
    nTp = plan.nT + 2 * int(plan.padding[1])
    s = _stream(plan.device)
    rad = plan.radon(x, flags=flags)
    dC = torch.empty((N, nRp, nTp), dtype=torch.float32, device=plan.device)
    _hip.check(lib.ebsd_band_grad(plan.handle, rad.data_ptr(), N, maxloc.data_ptr(), aveloc.data_ptr(),
                                  valid.data_ptr(), g_rho.data_ptr(), g_theta.data_ptr(), dC.data_ptr(), s))
    dR = conv_transpose(plan, dC)
    return backproject(plan, dR)


def conv_transpose(plan, dC):
    """Adjoint of padding + convolution: (N, nRp, nTp) -> (N, nR, nT)."""
    import torch
    dC = dC.to(torch.float32).contiguous()
    dR = torch.empty((dC.shape[0], plan.nR, plan.nT), dtype=torch.float32, device=plan.device)
    _hip.check(plan.lib.ebsd_conv_transpose(plan.handle, dC.data_ptr(), dC.shape[0], dR.data_ptr(),
                                            _stream(plan.device)))
    return dR


def backproject(plan, dR):
    """Adjoint of the normalised Radon: (N, nR, nT) -> (N, H, W)."""
    import torch
    dR = dR.to(torch.float32).contiguous()
    dX = torch.empty((dR.shape[0], plan.H, plan.W), dtype=torch.float32, device=plan.device)
    _hip.check(plan.lib.ebsd_backproject(plan.handle, dR.data_ptr(), dR.shape[0], dX.data_ptr(),
                                         _stream(plan.device)))
    return dX


def gnomonic_bands(normals, pc, H, W, vendor="EDAX"):
    """Predicted band (rho, theta) (N, M) float64 and d rho / d PC (N, M, 3) of
    detector-frame plane normals (N, M, 3) at pattern centre pc (3,) or (N, 3)."""
    import torch
    lib = _hip.load()
    n = normals.to(torch.float64).contiguous()
    N, M = n.shape[0], n.shape[1]
    P = torch.as_tensor(pc, dtype=torch.float64, device=n.device).contiguous()
    per = int(P.dim() == 2)
    rho = torch.empty((N, M), dtype=torch.float64, device=n.device)
    th = torch.empty_like(rho)
    J = torch.empty((N, M, 3), dtype=torch.float64, device=n.device)
    _hip.check(lib.ebsd_gnomonic_bands(N, M, n.data_ptr(), P.data_ptr(), per, int(H), int(W), _vendor(vendor),
                                       rho.data_ptr(), th.data_ptr(), J.data_ptr(), _stream(n.device)))
    return rho, th, J


def _gnomonic_fn():
    import torch

    class Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, normals, pc, H, W, vendor):
            rho, th, J = gnomonic_bands(normals.detach(), pc.detach(), H, W, vendor)
            ctx.save_for_backward(J)
            ctx.pc_shape = pc.shape
            ctx.mark_non_differentiable(th)
            return rho, th

        @staticmethod
        def backward(ctx, g_rho, g_theta):
            (J,) = ctx.saved_tensors
            if g_rho is None:
                return None, None, None, None, None
            g = (g_rho.to(torch.float64).unsqueeze(-1) * J)
            dpc = g.sum(dim=1) if len(ctx.pc_shape) == 2 else g.sum(dim=(0, 1))
            return None, dpc.reshape(ctx.pc_shape), None, None, None

    return Fn


class GnomonicBands:
    """``GnomonicBands.apply(normals, pc, H, W, vendor)`` -> (rho_hat, theta_hat);
    differentiable w.r.t. ``pc`` (a float64 tensor of shape (3,) or (N, 3))."""

    @staticmethod
    def apply(normals, pc, H, W, vendor="EDAX"):
        return _GnomonicFn.apply(normals, pc, H, W, vendor)


def radon2pole(theta, rho, pc, H, W, vendor="EDAX", jacobian=False):
    """``Radon.radon2pole`` on the device: (N, nB) band theta, rho -> (N, nB, 3) float32
    unit normals (and (N, nB, 3, 3) d n / d PC when ``jacobian``)."""
    import torch
    lib = _hip.load()
    th = torch.as_tensor(theta).to(torch.float32).contiguous()
    dev = th.device
    rh = torch.as_tensor(rho, device=dev).to(torch.float32).contiguous()
    P = torch.as_tensor(pc, dtype=torch.float64, device=dev).contiguous()
    N, nB = th.shape
    out = torch.empty((N, nB, 3), dtype=torch.float32, device=dev)
    J = torch.empty((N, nB, 3, 3), dtype=torch.float64, device=dev) if jacobian else None
    _hip.check(lib.ebsd_radon2pole(N, nB, th.data_ptr(), rh.data_ptr(), P.data_ptr(), int(P.dim() == 2), int(H),
                                   int(W), _vendor(vendor), out.data_ptr(), J.data_ptr() if jacobian else None,
                                   _stream(dev)))
    return (out, J) if jacobian else out


def match_bands(rho, theta, valid, rho_hat, theta_hat, rho_scale, max_dist=4.0):
    """Nearest predicted band for each valid detected band, in (rho, rho_scale * theta)
    (``rho_scale`` in pixels per radian, e.g. the largest |rho|, turns angles into detector
    distances) with the equivalence (rho, theta) ~ (-rho, theta - pi).  Returns (index
    (N, nB) int64, sign (N, nB) +-1 of the predicted rho, mask (N, nB) bool of matches
    within ``max_dist`` pixels)."""
    import torch
    rs = float(rho_scale)
    r = rho.to(torch.float64).unsqueeze(-1)
    t = theta.to(torch.float64).unsqueeze(-1)
    rh = rho_hat.unsqueeze(1)
    th = theta_hat.unsqueeze(1)
    d1 = (r - rh) ** 2 + (rs * (t - th)) ** 2
    dt = t - th
    alt = torch.where(dt > 0, th + np.pi, th - np.pi)
    d2 = (r + rh) ** 2 + (rs * (t - alt)) ** 2
    d = torch.minimum(d1, d2)
    best = d.argmin(dim=-1)
    dmin = d.gather(-1, best.unsqueeze(-1)).squeeze(-1)
    sgn = torch.where(d1.gather(-1, best.unsqueeze(-1)).squeeze(-1) <= d2.gather(-1, best.unsqueeze(-1)).squeeze(-1),
                      1.0, -1.0).to(torch.float64)
    mask = (valid > 0) & (dmin <= max_dist ** 2)
    return best, sgn, mask


def pc_loss(rho, theta, normals, pc, H, W, index, sign, mask, vendor="EDAX", rho_scale=1.0):
    """Squared error of detected band (rho, theta) against the bands predicted from the
    known orientations (normals) at ``pc``; differentiable w.r.t. ``pc``."""
    import torch
    rho_hat, theta_hat = GnomonicBands.apply(normals, pc, H, W, vendor)
    rh = sign * rho_hat.gather(1, index)
    th = theta_hat.gather(1, index)
    th = torch.where(sign < 0, torch.where(theta.to(torch.float64) > th, th + np.pi, th - np.pi), th)
    w = mask.to(torch.float64)
    dr = (rho.to(torch.float64) - rh) * w
    dt = (theta.to(torch.float64) - th) * w * rho_scale
    return (dr * dr).sum() + (dt * dt).sum()


def refine_pc(bd, pats, normals, pc0, vendor="EDAX", iters=8, tol=1e-9, max_dist=4.0):
    """Refine a shared pattern centre from ``pc0`` for patterns with known orientations.

    ``normals``: detector-frame plane normals (N, M, 3) of the bands those orientations
    predict.  Bands are detected once (``BandPeaks``), matched to the prediction at
    ``pc0``, then Gauss-Newton steps use the loss gradient and the d rho / d PC Jacobian
    of ``ebsd_gnomonic_bands``.  Returns (pc, history of (pc, loss))."""
    import torch
    H, W = int(bd.patDim[-2]), int(bd.patDim[-1])
    with torch.no_grad():
        rho, theta, valid = BandPeaks.apply(pats, bd)
    pc = torch.as_tensor(pc0, dtype=torch.float64, device=pats.device).clone()
    nrm = normals.to(device=pats.device, dtype=torch.float64)
    rho_hat, theta_hat, _ = gnomonic_bands(nrm, pc, H, W, vendor)
    idx, sgn, mask = match_bands(rho, theta, valid, rho_hat, theta_hat, float(bd.rhoMax), max_dist=max_dist)
    hist = []
    for _ in range(iters):
        p = pc.clone().requires_grad_(True)
        L = pc_loss(rho, theta, nrm, p, H, W, idx, sgn, mask, vendor)
        (g,) = torch.autograd.grad(L, p)
        # Gauss-Newton: J^T J from the per-band d rho / d PC of the matched bands
        _, _, J = gnomonic_bands(nrm, pc, H, W, vendor)
        Jm = (sgn.unsqueeze(-1) * J.gather(1, idx.unsqueeze(-1).expand(-1, -1, 3))) * mask.unsqueeze(-1)
        A = 2.0 * torch.einsum("nbi,nbj->ij", Jm, Jm)
        step = torch.linalg.solve(A + 1e-12 * torch.eye(3, dtype=A.dtype, device=A.device), g)
        hist.append((pc.detach().cpu().numpy().copy(), float(L.detach())))
        pc = (pc - step).detach()
        if float(step.abs().max()) < tol:
            break
    with torch.no_grad():
        L = pc_loss(rho, theta, nrm, pc, H, W, idx, sgn, mask, vendor)
    hist.append((pc.cpu().numpy().copy(), float(L)))
    return pc, hist


_BandPeaksFn = None
_GnomonicFn = None


def _init():
    global _BandPeaksFn, _GnomonicFn
    _BandPeaksFn = _band_peaks_fn()
    _GnomonicFn = _gnomonic_fn()


try:
    import torch as _torch  # noqa: F401
    _init()
except ImportError:  # torch is the device plumbing; without it there is nothing to run
    pass
