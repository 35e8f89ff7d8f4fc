"""Seeded synthetic cost volumes for the benchmark and test workloads.

Arrays are float32 in MATLAB layout: numpy shape (Ny, Nx[, Nz][, L]) in Fortran order,
so ``a.ravel(order="F")`` is the flat column-major buffer the solvers take.

binary2d_ellipses  256x256 image: union of 5 random ellipses, I ~ N(0.7, 0.1) inside and
                   N(0.3, 0.1) outside; Cs = |I - 0.7|, Ct = |I - 0.3|, alpha = 0.5.
                   pars [256; 256; 300; 0; 0.3; 0.16].
binary3d_ellipsoids  union of 12 random rotated ellipsoids (semi-axes U[10, 50]),
                   I ~ N(0.65, 0.12) inside and N(0.35, 0.12) outside; Gaussian negative
                   log-likelihoods Ct = (I-0.65)^2/(2*0.12^2) (foreground) and
                   Cs = (I-0.35)^2/(2*0.12^2) (background), so u ~ 1 on the foreground;
                   contrast-sensitive alpha = 0.2 + 0.8*exp(-20*|grad I|^2) with central
                   differences (one-sided at the border).  cc = 0.35, steps = 0.11.
potts3d_voronoi    L labels from L uniform random seeds: Voronoi indicator per label,
                   Gaussian-smoothed (sigma = 3 voxels), arg-max label map;
                   I ~ N(l/5, 0.08); Ct_l = (I - l/5)^2/(2*0.08^2); alpha = 0.8.
                   cc = 0.25, steps = 0.11.
ishikawa3d_voronoi the Potts volume above (labels ordered by intensity l/5, so the
                   linearly ordered model applies) with alpha_l = 0.3/(l+1) for the L-1 cut
                   functions, as tests/runTestIshikawa.m:119-121; cc = 0.75, steps = 0.1
                   (runTestIshikawa.m:127).
"""
from __future__ import annotations

import numpy as np


def _rot2(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


def binary2d_ellipses(n: int = 256, seed: int = 1001):
    rng = np.random.default_rng(seed)
    y, x = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64),
                       indexing="ij")
    fg = np.zeros((n, n), dtype=bool)
    for _ in range(5):
        cy, cx = rng.uniform(0, n, 2)
        ay, ax = rng.uniform(15, 60, 2)
        th = rng.uniform(0, np.pi)
        R = _rot2(th)
        dy, dx = y - cy, x - cx
        py = R[0, 0] * dy + R[1, 0] * dx
        px = R[0, 1] * dy + R[1, 1] * dx
        fg |= (py / ay) ** 2 + (px / ax) ** 2 <= 1.0
    I = np.where(fg, rng.normal(0.7, 0.1, (n, n)), rng.normal(0.3, 0.1, (n, n)))
    I = I.astype(np.float32)
    Cs = np.abs(I - np.float32(0.7)).astype(np.float32)
    Ct = np.abs(I - np.float32(0.3)).astype(np.float32)
    alpha = np.full((n, n), 0.5, dtype=np.float32)
    pars = np.array([n, n, 300, 0.0, 0.3, 0.16], dtype=np.float32)
    return (np.asfortranarray(Cs), np.asfortranarray(Ct), np.asfortranarray(alpha), pars, fg)


def _random_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _ellipsoid_mask(shape, rng, count=12, amin=10.0, amax=50.0):
    ny, nx, nz = shape
    fg = np.zeros(shape, dtype=bool)
    for _ in range(count):
        c = rng.uniform([0, 0, 0], [ny, nx, nz])
        a = rng.uniform(amin, amax, 3)
        R = _random_rotation(rng)
        r = int(np.ceil(a.max())) + 1
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, [ny, nx, nz])
        if np.any(hi <= lo):
            continue
        yy, xx, zz = np.meshgrid(np.arange(lo[0], hi[0]), np.arange(lo[1], hi[1]),
                                 np.arange(lo[2], hi[2]), indexing="ij")
        d = np.stack([yy - c[0], xx - c[1], zz - c[2]], axis=-1)
        p = d @ R  # coordinates in the ellipsoid frame
        inside = ((p / a) ** 2).sum(-1) <= 1.0
        fg[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= inside
    return fg


def _grad_sq(I):
    g = np.zeros_like(I, dtype=np.float32)
    for ax in range(I.ndim):
        d = np.gradient(I, axis=ax)  # central differences, one-sided at the borders
        g += d * d
    return g


def binary3d_ellipsoids(shape=(256, 256, 256), seed: int = 1002, count: int = 12,
                        return_image: bool = False):
    """Returns (Cs, Ct, alpha, fg) for the 3D binary workload (and the image I last when
    return_image)."""
    rng = np.random.default_rng(seed)
    fg = _ellipsoid_mask(tuple(shape), rng, count=count)
    I = rng.normal(0.0, 0.12, size=shape).astype(np.float32)
    I += np.where(fg, np.float32(0.65), np.float32(0.35))
    s2 = np.float32(2 * 0.12 ** 2)
    Ct = ((I - np.float32(0.65)) ** 2 / s2).astype(np.float32)
    Cs = ((I - np.float32(0.35)) ** 2 / s2).astype(np.float32)
    alpha = (np.float32(0.2) + np.float32(0.8) * np.exp(np.float32(-20.0) * _grad_sq(I)))
    alpha = alpha.astype(np.float32)
    out = ((np.asfortranarray(Cs), np.asfortranarray(Ct), np.asfortranarray(alpha), fg))
    return out + (I,) if return_image else out


BINARY3D_CC, BINARY3D_STEPS = 0.35, 0.11


def potts3d_voronoi(shape=(192, 192, 192), nlab: int = 6, seed: int = 1003):
    """Returns (Ct, alpha, labels) with Ct, alpha of shape (Ny, Nx, Nz, L)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    shape = tuple(shape)
    seeds = rng.uniform([0, 0, 0], shape, size=(nlab, 3))
    yy, xx, zz = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in shape], indexing="ij")
    d2 = np.stack([(yy - s[0]) ** 2 + (xx - s[1]) ** 2 + (zz - s[2]) ** 2 for s in seeds])
    vor = np.argmin(d2, axis=0)
    del d2
    sm = np.stack([gaussian_filter((vor == l).astype(np.float32), 3.0) for l in range(nlab)])
    lab = np.argmax(sm, axis=0)
    del sm
    mu = (np.arange(nlab, dtype=np.float32) / np.float32(5.0))
    I = (mu[lab] + rng.normal(0.0, 0.08, size=shape)).astype(np.float32)
    Ct = np.empty(shape + (nlab,), dtype=np.float32, order="F")
    for l in range(nlab):
        Ct[..., l] = (I - mu[l]) ** 2 / np.float32(2 * 0.08 ** 2)
    alpha = np.full(shape + (nlab,), 0.8, dtype=np.float32, order="F")
    return Ct, alpha, lab


POTTS3D_CC, POTTS3D_STEPS = 0.25, 0.11


def random_smooth_binary(shape, seed, alpha=0.25):
    """Smoothed random costs in the manner of tests/runTestBinaryMF.m:36-44 (per-slice
    Gaussian filtering of uniform noise), seeded."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    sig = (1.0, 1.0) + (0.0,) * (len(shape) - 2)
    Cs = gaussian_filter(rng.random(shape), sig).astype(np.float32)
    Ct = gaussian_filter(rng.random(shape), sig).astype(np.float32)
    a = np.full(shape, alpha, dtype=np.float32)
    return np.asfortranarray(Cs), np.asfortranarray(Ct), np.asfortranarray(a)


def random_smooth_potts(shape, nlab, seed, alpha_step=0.025):
    """Smoothed random label costs and per-label alpha = alpha_step*(l+1), in the manner of
    tests/runTestPotts.m:100-113, seeded."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    sig = (1.0, 1.0) + (0.0,) * (len(shape) - 2)
    Ct = np.empty(tuple(shape) + (nlab,), dtype=np.float32, order="F")
    alpha = np.empty(tuple(shape) + (nlab,), dtype=np.float32, order="F")
    for l in range(nlab):
        Ct[..., l] = gaussian_filter(rng.random(shape), sig)
        alpha[..., l] = alpha_step * (l + 1)
    return Ct, alpha


def ishikawa3d_voronoi(shape=(192, 192, 192), nlab: int = 6, seed: int = 1003):
    """Returns (Ct, alpha, labels): Ct (Ny, Nx, Nz, L), alpha (Ny, Nx, Nz, L-1)."""
    Ct, _, lab = potts3d_voronoi(shape, nlab, seed)
    alpha = np.empty(tuple(shape) + (nlab - 1,), dtype=np.float32, order="F")
    for l in range(nlab - 1):
        alpha[..., l] = np.float32(0.3 / (l + 1))
    return Ct, alpha, lab


ISHIKAWA3D_CC, ISHIKAWA3D_STEPS = 0.75, 0.1


def hmf_linear2d(shape=(256, 256), levels: int = 40, seed: int = 1005, noise: float = 0.15):
    """The linear (Ishikawa-like) reconstruction model of tests/runTestHMFLinearModel.m:11-27
    on a synthetic image (the test's cell.tif is not in the repository): a smooth image in
    [0, 1] plus Gaussian noise, leaf i of a flat tree with Ct_i = sqrt(|image - i/(L-1)|).
    Returns (costs [L arrays], noisy image, clean image)."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    clean = gaussian_filter(rng.random(shape), 6.0)
    clean = (clean - clean.min()) / (clean.max() - clean.min())
    img = clean + noise * rng.standard_normal(shape)
    costs = [np.sqrt(np.abs(img - i / (levels - 1))).astype(np.float32) for i in range(levels)]
    return costs, img.astype(np.float32), clean.astype(np.float32)


HMF2D_ALPHA, HMF2D_ITERS, HMF2D_STEPS = 0.15, 200, 0.1


def hmf3d_tree(shape=(192, 192, 192), seed: int = 1006):
    """A two-level label tree on the Voronoi intensity volume of potts3d_voronoi (6 labels,
    intensities l/5): root -> [dark -> (0, 1, 2), bright -> (3, 4, 5)].  Returns
    (parent, Ct per node (None for internal nodes), alpha per node, labels)."""
    Ct, _, lab = potts3d_voronoi(shape, nlab=6, seed=seed)
    parent = [-1, 0, 0, 1, 1, 1, 2, 2, 2]
    costs = [None, None, None] + [np.asfortranarray(Ct[..., k]) for k in range(6)]
    alpha = [0.0, 0.5, 0.5] + [1.0] * 6
    return parent, costs, alpha, lab


HMF3D_ITERS, HMF3D_CC, HMF3D_STEPS = 100, 0.25, 0.11
