"""Data costs, regularisation weights and discrete labels on the GPU (the steps the
reference's applications and tests run around the solvers).  Device tensors in, device
tensors out; every computation is a kernel of libasets_hip.so (include/asets.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat
from .solver import _ndim_dims, _ptr, _stream_handle

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

BINARY, POTTS, ISHIKAWA = 0, 1, 2


def _opts(stream):
    return nat.make_opts(_stream_handle(stream), nat.FLOW_F32, nat.KERNEL_AUTO, 0, 0, 0, None)


def cost_loglik(img, scribbles, nlab: int, epsilon: float = 1e-10, *, stream=None):
    """Ct [nlab * n] from an image and int32 scribbles (label l marks voxels == l),
    as computeLogLikelihoodCost of interactiveGraphCutSegmentation3D.m:96-118."""
    if scribbles.dtype != torch.int32 or scribbles.numel() != img.numel():
        raise ValueError("scribbles must be int32 of the image's size")
    n = img.numel()
    Ct = torch.empty(n * nlab, dtype=torch.float32, device=img.device)
    nat.check(nat.lib().asets_cost_loglik_dev(_ptr(img), scribbles.data_ptr(), n, nlab, epsilon,
                                              _ptr(Ct), C.byref(_opts(stream))))
    return Ct


def cost_neglog(p, epsilon: float = 1e-10, *, stream=None):
    """-log(p + epsilon) (atlasProbabilityMapsRegularization3D.m:47-53)."""
    Ct = torch.empty_like(p)
    nat.check(nat.lib().asets_cost_neglog_dev(_ptr(p), p.numel(), epsilon, _ptr(Ct),
                                              C.byref(_opts(stream))))
    return Ct


def cost_l1(img, models, *, stream=None):
    """Ct [len(models) * n]: |x - m_l| with x the image normalised to [0, 1]
    (L1IntensitySegmentation2D.m:36-44)."""
    m = (C.c_float * len(models))(*[float(v) for v in models])
    n = img.numel()
    Ct = torch.empty(n * len(models), dtype=torch.float32, device=img.device)
    nat.check(nat.lib().asets_cost_l1_dev(_ptr(img), n, m, len(models), _ptr(Ct),
                                          C.byref(_opts(stream))))
    return Ct


def fspecial_gaussian(size: int = 3, sigma: float = 0.5):
    """MATLAB fspecial('gaussian', [size size], sigma): exp(-(x^2+y^2)/(2 sigma^2)) on the
    centred grid, entries below eps*max set to 0, normalised to sum 1 (FP64)."""
    r = (size - 1) / 2.0
    x = np.arange(size) - r
    X, Y = np.meshgrid(x, x)
    h = np.exp(-(X * X + Y * Y) / (2.0 * sigma * sigma))
    h[h < np.finfo(np.float64).eps * h.max()] = 0.0
    s = h.ravel(order="F").sum()
    return h / s if s != 0 else h


def cost_smooth_labels(labels, shape, ids, h, *, stream=None):
    """Ct [len(ids) * n]: imfilter(1 - (labels == id), h, 'replicate') on a 2D int32 label
    image of shape (Ny, Nx) (smoothManualSegmentations2D.m:48-55)."""
    ny, nx = int(shape[0]), int(shape[1])
    if labels.dtype != torch.int32 or labels.numel() != ny * nx:
        raise ValueError("labels must be int32 of shape (Ny, Nx)")
    h = np.asarray(h, np.float64)
    if h.ndim != 2 or h.shape[0] != h.shape[1]:
        raise ValueError("h must be a square kernel")
    hf = np.ascontiguousarray(h.ravel(order="F"))
    idv = (C.c_int32 * len(ids))(*[int(v) for v in ids])
    Ct = torch.empty(ny * nx * len(ids), dtype=torch.float32, device=labels.device)
    nat.check(nat.lib().asets_cost_smooth_labels_dev(labels.data_ptr(), ny, nx, idv, len(ids),
                                                     hf.ctypes.data, h.shape[0], _ptr(Ct),
                                                     C.byref(_opts(stream))))
    return Ct


def alpha_contrast(img, shape, a0: float = 0.0, a1: float = 0.2, beta: float = 10.0,
                   power: int = 1, normalize: bool = True, *, stream=None):
    """a0 + a1*exp(-beta*|grad I|^power) with MATLAB gradient() (smoothManualSegmentations2D.m
    :73-82 uses the defaults)."""
    ndim, ny, nx, nz = _ndim_dims(shape)
    alpha = torch.empty_like(img)
    nat.check(nat.lib().asets_alpha_contrast_dev(ndim, _ptr(img), ny, nx, nz, a0, a1, beta, power,
                                                 int(normalize), _ptr(alpha),
                                                 C.byref(_opts(stream))))
    return alpha


def labels(model: int, u, n: int, nlab: int = 1, *, stream=None):
    """uint8 labels of n voxels: BINARY u > 0.5; POTTS 1-based first arg-max of nlab
    fields; ISHIKAWA number of the nlab-1 cut functions above 0.5."""
    lab = torch.empty(n, dtype=torch.uint8, device=u.device)
    nat.check(nat.lib().asets_labels_dev(model, _ptr(u), n, nlab, lab.data_ptr(),
                                         C.byref(_opts(stream))))
    return lab
