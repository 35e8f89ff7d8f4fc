"""CPU oracle of the cost, weight and label steps (TEST INFRASTRUCTURE ONLY: the checker
for asets_cost_loglik_dev, asets_cost_neglog_dev, asets_cost_l1_dev,
asets_cost_smooth_labels_dev, asets_alpha_contrast_dev and asets_labels_dev).  numpy
restatements of the reference's MATLAB:

  loglik   applications/interactiveGraphCutSegmentation3D.m:96-118
  neglog   applications/atlasProbabilityMapsRegularization3D.m:47-53
  l1       applications/L1IntensitySegmentation2D.m:36-44
  smooth   applications/smoothManualSegmentations2D.m:48-55 (imfilter, 'replicate')
  alpha    applications/smoothManualSegmentations2D.m:73-82 (MATLAB gradient())
  labels   tests/runTestBinaryMF.m:55, tests/runTestPotts.m:58,
           tests/runTestIshikawa.m:57-60
Arrays are flat in column-major order (shape (Ny, Nx[, Nz])).
"""
from __future__ import annotations

import numpy as np


def cost_loglik(img, scribbles, nlab, epsilon=1e-10):
    img = np.asarray(img, np.float32).ravel(order="F")
    scr = np.asarray(scribbles).ravel(order="F")
    mn, mx = img.min(), img.max()
    x = ((img - mn) / np.float32(mx - mn)) * np.float32(255.0)      # single, as img_n
    hist_bin = np.minimum(np.floor(x).astype(np.int64), 255)          # histc, edges 0..255
    look = np.floor(x).astype(np.int64)
    look += (x - look.astype(np.float32)) >= np.float32(0.5)          # uint16(): half away
    out = np.empty((nlab, img.size), np.float32)
    for l in range(1, nlab + 1):
        counts = np.bincount(hist_bin[scr == l], minlength=256).astype(np.float64)
        P = counts / counts.sum()
        out[l - 1] = (-np.log10(P[look] + epsilon)).astype(np.float32)
    return out.ravel()


def cost_neglog(p, epsilon=1e-10):
    return (-np.log(np.asarray(p, np.float64) + epsilon)).astype(np.float32)


def cost_l1(img, models):
    """|x - m_l| with x = (img - min) / (max - min), fp32 operations."""
    img = np.asarray(img, np.float32).ravel(order="F")
    x = (img - img.min()) / (img.max() - img.min())
    return np.concatenate([np.abs(x - np.float32(m)) for m in models]).astype(np.float32)


def cost_smooth_labels(lab, ids, h):
    """imfilter(1 - (lab == id), h, 'replicate') per id, on a 2D label image: the
    correlation sum over the K x K window written out, FP64, h's column-major order."""
    lab = np.asarray(lab)
    h = np.asarray(h, np.float64)
    K = h.shape[0]
    r = K // 2
    ny, nx = lab.shape
    pad = np.pad(lab, r, mode="edge")                 # 'replicate' border
    out = []
    for idv in ids:
        acc = np.zeros((ny, nx))
        for b in range(K):                            # column-major: a (dim 1) fastest
            for a in range(K):
                acc = acc + h[a, b] * (pad[a:a + ny, b:b + nx] != idv)
        out.append(acc.astype(np.float32).ravel(order="F"))
    return np.concatenate(out)


def matlab_gradient(I, axis):
    """numpy restatement of MATLAB gradient() along one axis (unit spacing)."""
    I = np.asarray(I, np.float64)
    g = np.zeros_like(I)
    n = I.shape[axis]
    if n < 2:
        return g
    sl = lambda a, b: tuple(slice(a, b) if d == axis else slice(None) for d in range(I.ndim))
    g[sl(0, 1)] = I[sl(1, 2)] - I[sl(0, 1)]
    g[sl(n - 1, n)] = I[sl(n - 1, n)] - I[sl(n - 2, n - 1)]
    if n > 2:
        g[sl(1, n - 1)] = (I[sl(2, n)] - I[sl(0, n - 2)]) * 0.5
    return g


def alpha_contrast(img, a0=0.0, a1=0.2, beta=10.0, power=1, normalize=True):
    """img (Ny, Nx[, Nz]); MATLAB's [gx, gy] = gradient(img): gx along columns (axis 1),
    gy along rows (axis 0), gz along slices."""
    I = np.asarray(img, np.float32).astype(np.float64)
    if normalize:
        mn, mx = float(np.float32(I.min())), float(np.float32(I.max()))
        I = (I - mn) * (1.0 / (mx - mn))
    q = sum(matlab_gradient(I, ax) ** 2 for ax in range(I.ndim))
    m = np.sqrt(q) if power == 1 else q
    return (a0 + a1 * np.exp(-beta * m)).astype(np.float32)


def labels(model, u, n, nlab=1):
    u = np.asarray(u, np.float32).reshape(-1)
    if model == 0:
        return (u[:n] > 0.5).astype(np.uint8)
    if model == 1:
        return (np.argmax(u[:nlab * n].reshape(nlab, n), axis=0) + 1).astype(np.uint8)
    return (u[:(nlab - 1) * n].reshape(nlab - 1, n) > 0.5).sum(0).astype(np.uint8)
