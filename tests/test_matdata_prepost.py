"""MAT v5 loading (asetsmatlabmaxflow_amd.matdata), the committed real-data fixtures, and
the pins of the oracle's L1-cost and smoothed-label-cost restatements (CPU only)."""
from pathlib import Path

import numpy as np
import pytest

import prepost_oracle as po
from asetsmatlabmaxflow_amd import matdata
from asetsmatlabmaxflow_amd.prepost import fspecial_gaussian

GOLDEN = Path(__file__).resolve().parent / "golden"
DATA = Path("/root/reference/data")


def test_load_mat_roundtrip(tmp_path):
    import scipy.io

    rng = np.random.default_rng(0)
    vals = {"img": rng.random((7, 5, 3)).astype(np.float32),
            "lab": rng.integers(0, 5, (7, 5)).astype(np.uint16),
            "d": rng.random((4,))}
    f = tmp_path / "t.mat"
    scipy.io.savemat(f, vals)
    got = matdata.load_mat(f)
    assert set(got) == set(vals)
    for k, v in vals.items():
        want = v.reshape(1, -1) if v.ndim == 1 else v      # MATLAB row vector
        assert got[k].dtype == want.dtype and got[k].shape == want.shape
        assert got[k].flags["F_CONTIGUOUS"]
        np.testing.assert_array_equal(got[k], want)
    assert set(matdata.load_mat(f, "lab")) == {"lab"}
    with pytest.raises(ValueError, match="no variable"):
        matdata.load_mat(f, ["nope"])
    scipy.io.savemat(tmp_path / "c.mat", {"c": np.array([[1, "x"]], dtype=object)})
    with pytest.raises(ValueError, match="not a numeric"):
        matdata.load_mat(tmp_path / "c.mat")
    assert ("img", (7, 5, 3), "single") in matdata.whos(f)


@pytest.mark.skipif(not DATA.exists(), reason="the reference's data files are not here")
def test_reference_data_files_and_fixture_agree():
    names = {n: (s, c) for n, s, c in matdata.whos(DATA / "brain_1125.mat")}
    assert names["img"] == ((91, 109, 91), "single")
    assert names["scribbles"] == ((91, 109, 91), "uint16")
    b = matdata.load_mat(DATA / "brain_1125.mat", ["img", "man_s40_flawed"])
    fx = np.load(GOLDEN / "data" / "brain_crops.npz")
    np.testing.assert_array_equal(fx["img3"], b["img"][0:91, 16:96, 41:49])
    np.testing.assert_array_equal(fx["man2"], b["man_s40_flawed"][:, :, 39])


def test_fixture_contents():
    fx = np.load(GOLDEN / "data" / "brain_crops.npz")
    assert fx["img3"].shape == (91, 80, 8) and fx["img3"].dtype == np.float32
    assert set(np.unique(fx["scribbles3"])) == {0, 1, 2, 3, 4}
    assert set(np.unique(fx["man2"])) == {0, 4}
    maps = [fx[k] for k in ("bg3", "csf3", "wm3", "gm3")]   # 8-bit quantised probabilities
    assert all(m.min() >= 0 and m.max() <= 1 for m in maps)
    s = sum(maps)
    assert s.min() > 1 - 1e-6 and s.max() < 1.2


def test_fspecial_gaussian_definition():
    for size, sigma in ((5, 1.5), (3, 0.5), (7, 2.0)):
        h = fspecial_gaussian(size, sigma)
        assert h.shape == (size, size)
        assert abs(h.sum() - 1) < 1e-15
        np.testing.assert_allclose(h, h.T, rtol=0, atol=0)
        np.testing.assert_allclose(h, h[::-1, ::-1], rtol=0, atol=1e-18)
        c = size // 2
        np.testing.assert_allclose(h[c, c + 1] / h[c, c], np.exp(-1 / (2 * sigma ** 2)),
                                   rtol=1e-14)


def test_l1_costs_closed_form():
    rng = np.random.default_rng(1)
    img = rng.uniform(3, 9, (13, 11)).astype(np.float32)
    models = (0.0, 0.25, 0.5, 0.9)
    Ct = po.cost_l1(img, models).reshape(4, -1)
    i_min, i_max = np.argmin(img.ravel(order="F")), np.argmax(img.ravel(order="F"))
    for l, m in enumerate(models):
        assert Ct[l, i_min] == np.float32(m)                  # x = 0
        assert Ct[l, i_max] == abs(np.float32(1) - np.float32(m))
    x = (img.ravel(order="F") - img.min()) / (img.max() - img.min())
    np.testing.assert_allclose(Ct[2], np.abs(x - 0.5), atol=1e-7)


def test_smooth_label_costs_pins():
    from scipy.ndimage import correlate

    rng = np.random.default_rng(2)
    h = fspecial_gaussian(5, 1.5)
    # constant image: 0 for its own label, the kernel's sum (1) for another
    lab = np.full((9, 12), 3)
    Ct = po.cost_smooth_labels(lab, [3, 7], h).reshape(2, 9, 12, order="C")
    assert not Ct[0].any()
    np.testing.assert_allclose(Ct[1], 1.0, atol=1e-7)
    # one marked pixel far from the border: 1 - h around it (h is symmetric)
    lab = np.zeros((21, 23), int)
    lab[10, 11] = 5
    c = po.cost_smooth_labels(lab, [5], h).reshape(23, 21).T
    want = np.ones((21, 23))
    want[8:13, 9:14] -= h
    np.testing.assert_allclose(c, want, atol=1e-7)
    # against an independent FP64 correlation with replicated borders
    lab = rng.integers(0, 3, (17, 15))
    got = po.cost_smooth_labels(lab, [0, 1, 2], h).reshape(3, 15, 17).transpose(0, 2, 1)
    for k in range(3):
        ref = correlate((lab != k).astype(np.float64), h, mode="nearest")
        np.testing.assert_allclose(got[k], ref, atol=2e-7)
    # the labels partition the image: sum over labels of (1 - Ct) is the kernel sum
    np.testing.assert_allclose((1 - got).sum(0), 1.0, atol=1e-6)
