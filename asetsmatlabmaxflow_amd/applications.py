"""The computational part of the reference's four application scripts, on the GPU.

Each function runs one script's pipeline end to end on device memory -- data costs and
weights, the Potts solve with the script's pars, the arg-max labelling -- through
libasets_hip.so; plotting is left out.  Inputs are arrays in MATLAB's shape (numpy or
torch), e.g. from matdata.load_mat of the reference's data files.

  interactive_segmentation_3d   applications/interactiveGraphCutSegmentation3D.m:31-50
  atlas_regularization_3d       applications/atlasProbabilityMapsRegularization3D.m:32-63
  smooth_manual_segmentation_2d applications/smoothManualSegmentations2D.m:31-85
  l1_intensity_segmentation_2d  applications/L1IntensitySegmentation2D.m:23-51
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import prepost, solver

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


@dataclass
class Segmentation:
    labels: "torch.Tensor"   # uint8, 1-based arg-max label per voxel (flat, column-major)
    u: "torch.Tensor"        # the solver's labelling functions [L * N]
    result: object           # solver.Result (cvg, it_num, seconds)
    shape: tuple
    nlab: int

    def label_image(self) -> np.ndarray:
        """Labels in MATLAB's shape (the index I of [uu, I] = max(u, [], ndims))."""
        return self.labels.cpu().numpy().reshape(self.shape, order="F")


def _dev(a, dtype=torch.float32 if torch else None, device="cuda"):
    """Column-major flat device copy of an array in MATLAB's shape."""
    if torch.is_tensor(a):
        t = a.to(device=device, dtype=dtype)
        return t.permute(*range(t.dim() - 1, -1, -1)).contiguous().reshape(-1)
    a = np.asarray(a)
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(device=device,
                                                                         dtype=dtype)


def _potts(Ct, alpha_field, shape, nlab, pars, stream=None):
    max_iter, err_bound, cc, steps = pars
    r = solver.potts(Ct, alpha_field, shape, nlab, max_iter, err_bound, cc, steps,
                     stream=stream)
    n = int(np.prod(shape))
    lab = prepost.labels(prepost.POTTS, r.u, n, nlab, stream=stream)
    return Segmentation(lab, r.u, r, tuple(shape), nlab)


def interactive_segmentation_3d(img, scribbles, *, alpha=0.025, epsilon=1e-10,
                                pars=(200, 1e-11, 0.25, 0.11), stream=None) -> Segmentation:
    """Log-likelihood costs from scribbles 1..L (computeLogLikelihoodCost), constant
    alpha, asetsPotts3D with pars [r c s L 200 1e-11 0.25 0.11], arg-max."""
    shape = tuple(np.shape(img))
    scr = np.asarray(scribbles.cpu() if torch.is_tensor(scribbles) else scribbles)
    ids = np.unique(scr[scr != 0])
    nlab = len(ids)
    if nlab == 0 or not np.array_equal(ids, np.arange(1, nlab + 1)):
        raise ValueError("scribbles must mark labels 1..L (the script tests scribbles == i)")
    Ct = prepost.cost_loglik(_dev(img), _dev(scr, torch.int32), nlab, epsilon, stream=stream)
    a = torch.full_like(Ct, float(alpha))
    return _potts(Ct, a, shape, nlab, pars, stream)


def atlas_regularization_3d(maps, *, alpha=0.05, epsilon=1e-10,
                            pars=(200, 1e-11, 0.25, 0.11), stream=None) -> Segmentation:
    """Ct_l = -log(p_l + eps) for the probability maps (bg, csf, wm, gm in the script),
    constant alpha, asetsPotts3D, arg-max."""
    shape = tuple(np.shape(maps[0]))
    Ct = torch.cat([prepost.cost_neglog(_dev(m), epsilon, stream=stream) for m in maps])
    a = torch.full_like(Ct, float(alpha))
    return _potts(Ct, a, shape, len(maps), pars, stream)


def smooth_manual_segmentation_2d(img, man, *, alpha=0.2, size=5, sigma=1.5,
                                  contrast=(0.0, 0.2, 10.0),
                                  pars=(300, 1e-11, 0.25, 0.11), stream=None):
    """Costs = Gaussian-smoothed inverse indicators of the manual labels (sorted unique
    values, 0 included), solved twice with asetsPotts2D: constant alpha, then the
    contrast-sensitive alpha a0 + a1*exp(-beta*|grad img|) of the normalised image.
    Returns (constant-alpha segmentation, contrast-alpha segmentation, label ids)."""
    shape = tuple(np.shape(man))
    m = np.asarray(man.cpu() if torch.is_tensor(man) else man)
    ids = [int(v) for v in np.sort(np.unique(m))]
    nlab = len(ids)
    h = prepost.fspecial_gaussian(size, sigma)
    Ct = prepost.cost_smooth_labels(_dev(m, torch.int32), shape, ids, h, stream=stream)
    a1 = torch.full_like(Ct, float(alpha))
    seg1 = _potts(Ct, a1, shape, nlab, pars, stream)
    a0, a_1, beta = contrast
    amap = prepost.alpha_contrast(_dev(img), shape, a0, a_1, beta, 1, True, stream=stream)
    seg2 = _potts(Ct, amap.repeat(nlab), shape, nlab, pars, stream)
    return seg1, seg2, ids


def l1_intensity_segmentation_2d(img, *, models=(0.0, 0.25, 0.5, 0.9), alpha=0.025,
                                 pars=(200, 1e-11, 0.25, 0.11), stream=None) -> Segmentation:
    """Ct_l = |img_normalised - m_l|, constant alpha, asetsPotts2D, arg-max."""
    shape = tuple(np.shape(img))
    Ct = prepost.cost_l1(_dev(img), models, stream=stream)
    a = torch.full_like(Ct, float(alpha))
    return _potts(Ct, a, shape, len(models), pars, stream)
