"""Seeded synthetic Kikuchi-like EBSD patterns (the workload generator for tests and bench).

Geometry is the exact inverse of ``Radon.radon2pole`` plus find_bands' theta/rho flip
(pyebsdindex/radon_fast.py:363-396, band_detect.py:519-524): pixel (row n, col m) of an
H x W pattern sits at detector point (m + 0.5, H - n - 0.5, 0), and the projection centre
is t = (x*.W, y*.H, -z*.min(H, W)) (EDAX convention).  A lattice plane with unit normal g
(detector frame) produces a band of intensity A exp(-a^2 / (2 s^2)), a = angle between the
ray t->pixel and the plane, s = half-width / sqrt(2 ln 2).  The background is a 2-D
Gaussian centred on the pattern with sigma = 0.6 * width; Poisson noise follows, then
clipping to the integer type.  Runs on any torch device.
"""
import itertools
import math

import numpy as np

FAMILIES = {
    "FCC": [(1, 1, 1), (2, 0, 0), (2, 2, 0), (3, 1, 1)],
    "BCC": [(1, 1, 0), (2, 0, 0), (2, 1, 1), (3, 1, 0)],
    # hexagonal (h k i l), c/a = 1.587 (alpha-Ti-like)
    "HCP": [(1, 0, -1, 0), (0, 0, 0, 2), (1, 0, -1, 1), (1, 0, -1, 2), (1, 1, -2, 0)],
}


def _dedupe(vecs):
    out = []
    for v in vecs:
        v = np.asarray(v, dtype=np.float64)
        v = v / np.linalg.norm(v)
        if not any(abs(abs(np.dot(v, u)) - 1.0) < 1e-9 for u in out):
            out.append(v)
    return np.array(out)


def plane_normals(family="FCC", c_over_a=1.587):
    """Unit plane normals (crystal frame) of all symmetric variants, one per +/- pair."""
    fam = family.upper()
    vecs = []
    if fam in ("FCC", "BCC"):
        for hkl in FAMILIES[fam]:
            for perm in set(itertools.permutations(hkl)):
                for sg in itertools.product((1, -1), repeat=3):
                    vecs.append(np.array(perm) * np.array(sg))
    elif fam == "HCP":
        a1 = np.array([1.0, 0.0, 0.0])
        a2 = np.array([-0.5, math.sqrt(3) / 2, 0.0])
        a3 = np.array([0.0, 0.0, c_over_a])
        V = np.dot(a1, np.cross(a2, a3))
        b1, b2, b3 = np.cross(a2, a3) / V, np.cross(a3, a1) / V, np.cross(a1, a2) / V
        for h, k, i, l in FAMILIES["HCP"]:
            for perm in set(itertools.permutations((h, k, i))):
                for s in (1, -1):
                    for sl in (1, -1):
                        hh, kk, _ = (s * perm[0], s * perm[1], s * perm[2])
                        vecs.append(hh * b1 + kk * b2 + sl * l * b3)
    else:
        raise ValueError(f"unknown family {family}")
    return _dedupe([v for v in vecs if np.linalg.norm(v) > 0])


def random_quaternions(n, rng):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def quat_to_matrix(q):
    """Rotation matrices (n,3,3) taking crystal-frame vectors to the detector frame."""
    q = np.asarray(q, dtype=np.float64)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], -2)


def kikuchi_patterns(n, shape=(60, 60), dtype="uint8", seed=0, family="FCC", pc=(0.5, 0.5, 0.6),
                     device="cpu", noise=True, background=None, amplitude=None,
                     halfwidth_deg=(1.0, 2.5), chunk=512, return_meta=False):
    """Generate ``n`` patterns as a torch tensor (n, H, W) of ``dtype`` on ``device``.

    Defaults follow the 8-bit workload (background peak 120, band amplitude 20-60); for
    uint16 the background peak is 2000 and the amplitude 300-900.  ``pc`` may be a
    3-vector or an (n, 3) array."""
    import torch
    H, W = int(shape[0]), int(shape[1])
    rng = np.random.default_rng(seed)
    dt = {"uint8": torch.uint8, "uint16": torch.uint16, "float32": torch.float32}[str(dtype)]
    if background is None:
        background = 2000.0 if dt == torch.uint16 else 120.0
    if amplitude is None:
        amplitude = (300.0, 900.0) if dt == torch.uint16 else (20.0, 60.0)
    normals = plane_normals(family)
    nb = normals.shape[0]
    quats = random_quaternions(n, rng)
    rot = quat_to_matrix(quats)
    g_det = np.einsum("nij,bj->nbi", rot, normals)                  # (n, nb, 3)
    amp = rng.uniform(amplitude[0], amplitude[1], size=(n, nb))
    hw = np.deg2rad(rng.uniform(halfwidth_deg[0], halfwidth_deg[1], size=(n, nb)))
    sig = hw / math.sqrt(2.0 * math.log(2.0))
    pcs = np.broadcast_to(np.asarray(pc, dtype=np.float64), (n, 3)).copy()
    t = np.stack([pcs[:, 0] * W, pcs[:, 1] * H, -pcs[:, 2] * min(H, W)], -1)

    dev = torch.device(device)
    m = torch.arange(W, dtype=torch.float64, device=dev) + 0.5
    y = H - torch.arange(H, dtype=torch.float64, device=dev) - 0.5
    X = m[None, :].expand(H, W)
    Y = y[:, None].expand(H, W)
    sb = 0.6 * W
    bgimg = background * torch.exp(-((torch.arange(W, device=dev, dtype=torch.float64)[None, :] - W / 2) ** 2 +
                                      (torch.arange(H, device=dev, dtype=torch.float64)[:, None] - H / 2) ** 2)
                                    / (2 * sb * sb))
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) + 7919)
    out = torch.empty((n, H, W), dtype=dt, device=dev)
    for s0 in range(0, n, chunk):
        s1 = min(n, s0 + chunk)
        tt = torch.as_tensor(t[s0:s1], device=dev)
        dx = X[None] - tt[:, 0, None, None]
        dy = Y[None] - tt[:, 1, None, None]
        dz = -tt[:, 2, None, None].expand(-1, H, W)
        dn = torch.sqrt(dx * dx + dy * dy + dz * dz)
        gg = torch.as_tensor(g_det[s0:s1], device=dev)
        dot = (gg[:, :, 0, None, None] * dx[:, None] + gg[:, :, 1, None, None] * dy[:, None] +
               gg[:, :, 2, None, None] * dz[:, None]) / dn[:, None]
        ang = torch.asin(torch.clamp(dot.abs(), max=1.0))
        ss = torch.as_tensor(sig[s0:s1], device=dev)[:, :, None, None]
        aa = torch.as_tensor(amp[s0:s1], device=dev)[:, :, None, None]
        img = (aa * torch.exp(-ang * ang / (2 * ss * ss))).sum(1) + bgimg[None]
        if noise and dt != torch.float32:
            img = torch.poisson(img, generator=gen)
        if dt == torch.uint8:
            img = img.round().clamp(0, 255)
        elif dt == torch.uint16:
            img = img.round().clamp(0, 65535)
        out[s0:s1] = img.to(dt)
    if return_meta:
        return out, {"quat": quats, "normals": normals, "g_det": g_det, "pc": pcs, "amp": amp,
                     "halfwidth": hw}
    return out


def grain_scan(nrows, ncols, shape=(60, 60), dtype="uint8", seed=0, grains=4, family="FCC", device="cpu"):
    """A scan (nrows*ncols, H, W) of ``grains`` orientations in vertical stripes: every
    scan point is its grain's noise-free pattern with its own Poisson noise (for integer
    types), the input NLPAR averages over."""
    import torch
    dt = {"uint8": torch.uint8, "uint16": torch.uint16, "float32": torch.float32}[str(dtype)]
    clean = kikuchi_patterns(grains, shape, "float32", seed=seed, family=family, device=device,
                             background=2000.0 if dt == torch.uint16 else 120.0,
                             amplitude=(300.0, 900.0) if dt == torch.uint16 else (20.0, 60.0))
    gid = (torch.arange(ncols, device=device) * grains // ncols).repeat(nrows)
    gen = torch.Generator(device=torch.device(device))
    gen.manual_seed(int(seed) + 104729)
    out = torch.empty((nrows * ncols,) + tuple(shape), dtype=dt, device=device)
    step = 4096
    for s0 in range(0, nrows * ncols, step):
        x = clean[gid[s0:s0 + step]].clamp(min=0).double()
        if dt != torch.float32:
            x = torch.poisson(x, generator=gen).clamp(0, 255 if dt == torch.uint8 else 65535)
        out[s0:s0 + step] = x.to(dt)
    return out


def cubic_band_normals(families, npats, nbands=9, seed=0, noise_deg=0.3, distinct=4096):
    """Band normals of a cubic phase as radon2pole would give them for npats patterns:
    ``distinct`` random orientations, each with nbands distinct poles (no two within 3
    degrees, antiparallel counted as one) drawn from ``families`` (list of (n_i, 3)
    integer pole arrays, e.g. BandIndexer.families), repeated to npats patterns, each
    tilted by Gaussian noise of noise_deg degrees per component.  float32 (npats, nbands,
    3) and the true family of each band (npats, nbands) int32."""
    rng = np.random.default_rng(seed)
    allp = np.concatenate([np.asarray(f, np.float64) / np.linalg.norm(f, axis=1, keepdims=True) for f in families])
    famid = np.concatenate([np.full(len(f), i) for i, f in enumerate(families)])
    cos3 = math.cos(math.radians(3.0))
    m = min(distinct, npats)
    base = np.zeros((m, nbands, 3))
    bfam = np.zeros((m, nbands), np.int32)
    rots = quat_to_matrix(random_quaternions(m, rng))
    for p in range(m):
        chosen = []
        for idx in rng.permutation(allp.shape[0]):
            if all(abs(allp[idx] @ allp[c]) < cos3 for c in chosen):
                chosen.append(idx)
                if len(chosen) == nbands:
                    break
        base[p, :len(chosen)] = allp[chosen] @ rots[p].T
        bfam[p, :len(chosen)] = famid[chosen]
    pick = np.arange(npats) % m
    n = base[pick] + rng.normal(0.0, math.radians(noise_deg), (npats, nbands, 3))
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    return n.astype(np.float32), bfam[pick]
