"""GPU: the L1 and smoothed-label cost kernels against the oracle, and the four
application pipelines (asetsmatlabmaxflow_amd.applications) on crops of the reference's
real data (tests/golden/data/brain_crops.npz) against the same pipeline composed of oracle
steps -- costs, the FP32 Potts solve, the arg-max -- bit for bit."""
from pathlib import Path

import numpy as np
import pytest

import oracle as orc
import prepost_oracle as po

pytestmark = pytest.mark.gpu
FX = Path(__file__).resolve().parent / "golden" / "data" / "brain_crops.npz"


def _fx():
    return np.load(FX)


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F")))
    return t.to("cuda", dtype=dtype or torch.float32)


def test_l1_and_smooth_label_costs_bitwise(cuda):
    import torch

    from asetsmatlabmaxflow_amd import prepost
    rng = np.random.default_rng(3)
    img = rng.uniform(-2, 5, (37, 29)).astype(np.float32)
    models = (0.0, 0.25, 0.5, 0.9, 0.333)
    got = prepost.cost_l1(_dev(img), models).cpu().numpy()
    np.testing.assert_array_equal(got, po.cost_l1(img, models))
    lab = rng.integers(0, 4, (37, 29)).astype(np.int32)
    for size, sigma in ((5, 1.5), (3, 0.5), (7, 2.5)):
        h = prepost.fspecial_gaussian(size, sigma)
        got = prepost.cost_smooth_labels(_dev(lab, torch.int32), lab.shape, [0, 1, 2, 3, 9],
                                         h).cpu().numpy()
        np.testing.assert_array_equal(got, po.cost_smooth_labels(lab, [0, 1, 2, 3, 9], h))


def _oracle_potts(Ct, alpha, shape, nlab, pars):
    it, eb, cc, st = pars
    r = orc.potts(Ct, alpha, shape, nlab, it, eb, cc, st)
    return r, po.labels(1, r.u, int(np.prod(shape)), nlab)


PARS = (40, 1e-11, 0.25, 0.11)   # the scripts' pars with fewer iterations


def test_interactive_segmentation_3d(cuda):
    from asetsmatlabmaxflow_amd import applications
    fx = _fx()
    img, scr = fx["img3"], fx["scribbles3"]
    seg = applications.interactive_segmentation_3d(img, scr, pars=PARS)
    Ct = po.cost_loglik(img, scr.astype(np.int32), 4, 1e-10)
    r, lab = _oracle_potts(Ct, np.full(Ct.shape, 0.025, np.float32), img.shape, 4, PARS)
    np.testing.assert_array_equal(seg.u.cpu().numpy(), r.u)
    np.testing.assert_array_equal(seg.labels.cpu().numpy(), lab)
    assert seg.label_image().shape == img.shape


def test_atlas_regularization_3d(cuda):
    from asetsmatlabmaxflow_amd import applications
    fx = _fx()
    maps = [fx[k] for k in ("bg3", "csf3", "wm3", "gm3")]
    seg = applications.atlas_regularization_3d(maps, pars=PARS)
    Ct = np.concatenate([po.cost_neglog(m.ravel(order="F")) for m in maps])
    r, lab = _oracle_potts(Ct, np.full(Ct.shape, 0.05, np.float32), maps[0].shape, 4, PARS)
    np.testing.assert_array_equal(seg.u.cpu().numpy(), r.u)
    np.testing.assert_array_equal(seg.labels.cpu().numpy(), lab)


def test_smooth_manual_segmentation_2d(cuda):
    from asetsmatlabmaxflow_amd import applications, prepost
    fx = _fx()
    img, man = fx["img2"], fx["man2"]
    seg1, seg2, ids = applications.smooth_manual_segmentation_2d(img, man)
    assert ids == [0, 4]
    h = prepost.fspecial_gaussian(5, 1.5)
    Ct = po.cost_smooth_labels(man, ids, h)
    pars = (300, 1e-11, 0.25, 0.11)
    r1, lab1 = _oracle_potts(Ct, np.full(Ct.shape, 0.2, np.float32), img.shape, 2, pars)
    np.testing.assert_array_equal(seg1.u.cpu().numpy(), r1.u)
    np.testing.assert_array_equal(seg1.labels.cpu().numpy(), lab1)
    a = po.alpha_contrast(img, 0.0, 0.2, 10.0, 1, True)
    r2, lab2 = _oracle_potts(Ct, np.tile(a.ravel(order="F"), 2), img.shape, 2, pars)
    np.testing.assert_array_equal(seg2.u.cpu().numpy(), r2.u)
    np.testing.assert_array_equal(seg2.labels.cpu().numpy(), lab2)


def test_l1_intensity_segmentation_2d(cuda):
    from asetsmatlabmaxflow_amd import applications
    fx = _fx()
    img = fx["img2"]
    seg = applications.l1_intensity_segmentation_2d(img)
    models = (0.0, 0.25, 0.5, 0.9)
    Ct = po.cost_l1(img, models)
    pars = (200, 1e-11, 0.25, 0.11)
    r, lab = _oracle_potts(Ct, np.full(Ct.shape, 0.025, np.float32), img.shape, 4, pars)
    np.testing.assert_array_equal(seg.u.cpu().numpy(), r.u)
    np.testing.assert_array_equal(seg.labels.cpu().numpy(), lab)
    # the labelling follows the intensity models: darker voxels, lower labels
    x = (img - img.min()) / (img.max() - img.min())
    L = seg.label_image()
    assert x[L == 1].mean() < x[L == 4].mean()
