"""Pins of the cost / weight / label restatements in oracle/prepost_oracle.py: closed
forms of the log-likelihood cost (including the reference's floor-binned histogram read
back at the rounded bin), -log of probabilities, MATLAB gradient() on polynomials, the
weight of a flat image, and label rules with ties."""
import numpy as np
import pytest

import prepost_oracle as PO


def test_loglik_two_classes_closed_form():
    img = np.array([0.0, 1.0, 2.0, 3.0], np.float32)        # x = 0, 85, 170, 255
    scr = np.array([1, 1, 2, 2], np.int32)
    Ct = PO.cost_loglik(img, scr, 2).reshape(2, 4)
    half = -np.log10(0.5 + 1e-10)
    np.testing.assert_allclose(Ct[0], [half, half, 10.0, 10.0], rtol=1e-6)
    np.testing.assert_allclose(Ct[1], [10.0, 10.0, half, half], rtol=1e-6)


def test_loglik_floor_binning_vs_rounded_lookup():
    """histc puts x = 84.6 in bin 84, the lookup uint16(84.6) reads bin 85: a scribbled
    voxel can get the epsilon cost of an empty bin (the reference's own behaviour)."""
    img = np.array([0.0, 84.6, 84.6, 255.0], np.float32)
    scr = np.array([0, 1, 1, 2], np.int32)
    Ct = PO.cost_loglik(img, scr, 2).reshape(2, 4)
    assert Ct[0, 1] == pytest.approx(10.0, rel=1e-6)         # bin 85 is empty for label 1
    assert Ct[1, 3] == pytest.approx(-np.log10(1.0 + 1e-10), abs=1e-7)   # x = 255: last bin


def test_neglog():
    p = np.array([1.0, np.exp(-2.0), 0.0], np.float32)
    np.testing.assert_allclose(PO.cost_neglog(p), [0.0, 2.0, -np.log(1e-10)], rtol=1e-6,
                               atol=1e-9)


def test_matlab_gradient_polynomials():
    n = 7
    x = np.arange(n, dtype=np.float64)
    ramp = np.tile((3.0 * x)[None, :], (4, 1))               # varies along axis 1
    np.testing.assert_array_equal(PO.matlab_gradient(ramp, 1), np.full((4, n), 3.0))
    np.testing.assert_array_equal(PO.matlab_gradient(ramp, 0), np.zeros((4, n)))
    sq = x ** 2
    g = PO.matlab_gradient(sq, 0)
    assert g[0] == 1.0 and g[-1] == 2 * n - 3
    np.testing.assert_array_equal(g[1:-1], 2 * x[1:-1])
    assert not PO.matlab_gradient(np.ones((1, 5)), 0).any()   # extent 1: zero gradient


def test_alpha_flat_and_ramp():
    flat = np.full((6, 5, 4), 7.0, np.float32)
    np.testing.assert_allclose(PO.alpha_contrast(flat, 0.1, 0.5, 10.0, 1, normalize=False),
                               0.6, rtol=1e-7)
    ramp = np.tile(np.arange(5, dtype=np.float32)[None, :], (6, 1))   # slope 1/4 normalised
    a = PO.alpha_contrast(ramp, 0.0, 0.2, 10.0, 1, normalize=True)
    np.testing.assert_allclose(a, 0.2 * np.exp(-10.0 * 0.25), rtol=1e-6)


def test_labels_rules():
    u = np.array([0.5, 0.51, 0.0, 1.0], np.float32)
    np.testing.assert_array_equal(PO.labels(0, u, 4), [0, 1, 0, 1])
    up = np.array([[0.2, 0.5, 0.4], [0.5, 0.5, 0.1], [0.3, 0.0, 0.5]], np.float32)  # [L, n]
    np.testing.assert_array_equal(PO.labels(1, up, 3, 3), [2, 1, 3])   # ties: first max
    ui = np.array([[0.9, 0.9, 0.2], [0.6, 0.4, 0.1]], np.float32)      # [L-1, n]
    np.testing.assert_array_equal(PO.labels(2, ui, 3, 3), [2, 1, 0])
