"""Pins of the FP64 energy evaluator of the oracle (orc_energy_binary / orc_energy_potts):
closed forms for constant fields, a straight edge and a single voxel, positive
homogeneity and shift invariance of the regulariser, the Nz = 1 == 2D identity and the
Potts-as-sum-of-fields identity."""
import numpy as np
import pytest

import oracle as orc

S2 = np.sqrt(2.0)


def test_constant_fields_give_the_data_term():
    rng = np.random.default_rng(0)
    sh = (5, 6, 7)
    Cs, Ct, a = rng.uniform(0, 3, sh), rng.uniform(0, 3, sh), rng.uniform(0.1, 1, sh)
    e0 = orc.energy_binary(Cs, Ct, a, np.zeros(sh), sh)
    e1 = orc.energy_binary(Cs, Ct, a, np.ones(sh), sh)
    assert e0[2] == 0 and e1[2] == 0
    assert e0[0] == pytest.approx(Cs.sum(), rel=1e-14)
    assert e1[0] == pytest.approx(Ct.sum(), rel=1e-14)
    # the data term is affine in u: u = 1/2 gives the mean of the two
    eh = orc.energy_binary(Cs, Ct, a, np.full(sh, 0.5), sh)
    assert eh[1] == pytest.approx(0.5 * (Cs.sum() + Ct.sum()), rel=1e-14)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_straight_edge_costs_sqrt2_alpha_per_face(axis):
    sh = (6, 7, 5)
    u = np.zeros(sh)
    idx = [slice(None)] * 3
    idx[axis] = slice(0, 3)
    u[tuple(idx)] = 1.0
    alpha = 0.7
    e = orc.energy_binary(np.zeros(sh), np.zeros(sh), np.full(sh, alpha), u, sh)
    faces = np.prod(sh) // sh[axis]
    assert e[2] == pytest.approx(S2 * alpha * faces, rel=1e-14)


def test_straight_edge_2d():
    sh = (6, 7)
    u = np.zeros(sh)
    u[:, :4] = 1.0
    e = orc.energy_binary(np.zeros(sh), np.zeros(sh), np.full(sh, 2.0), u, sh)
    assert e[2] == pytest.approx(S2 * 2.0 * 6, rel=1e-14)


def test_single_voxel():
    sh = (5, 5, 5)
    u = np.zeros(sh)
    u[2, 2, 2] = 1.0
    e = orc.energy_binary(np.zeros(sh), np.zeros(sh), np.full(sh, 1.5), u, sh)
    assert e[2] == pytest.approx(1.5 * (np.sqrt(3.0) + 6 / S2), rel=1e-14)
    sh2 = (5, 5)
    u2 = np.zeros(sh2)
    u2[2, 2] = 1.0
    e2 = orc.energy_binary(np.zeros(sh2), np.zeros(sh2), np.ones(sh2), u2, sh2)
    assert e2[2] == pytest.approx(np.sqrt(2.0) + 4 / S2, rel=1e-14)
    # a corner voxel: 3 faces inside the volume
    u[:] = 0
    u[0, 0, 0] = 1.0
    e = orc.energy_binary(np.zeros(sh), np.zeros(sh), np.ones(sh), u, sh)
    assert e[2] == pytest.approx(np.sqrt(1.5) + 3 / S2, rel=1e-14)


def test_regulariser_homogeneous_and_shift_invariant():
    rng = np.random.default_rng(1)
    sh = (7, 5, 6)
    u = rng.uniform(0, 1, sh)
    a = rng.uniform(0.2, 1, sh)
    z = np.zeros(sh)
    r = orc.energy_binary(z, z, a, u, sh)[2]
    assert orc.energy_binary(z, z, a, -2.5 * u, sh)[2] == pytest.approx(2.5 * r, rel=1e-13)
    assert orc.energy_binary(z, z, a, u + 3.0, sh)[2] == pytest.approx(r, rel=1e-12)
    # alpha enters linearly
    assert orc.energy_binary(z, z, 2 * a, u, sh)[2] == pytest.approx(2 * r, rel=1e-14)


def test_nz1_equals_2d():
    rng = np.random.default_rng(2)
    sh = (9, 8)
    Cs, Ct, a, u = (rng.uniform(0, 1, sh) for _ in range(4))
    e2 = orc.energy_binary(Cs, Ct, a, u, sh)
    e3 = orc.energy_binary(Cs, Ct, a, u, (9, 8, 1))
    np.testing.assert_allclose(e2, e3, rtol=1e-15)


def test_potts_is_sum_of_label_fields():
    rng = np.random.default_rng(3)
    sh, L = (6, 5, 4), 3
    Ct, a, u = (rng.uniform(0, 1, sh + (L,)) for _ in range(3))
    e = orc.energy_potts(Ct, a, u, sh, L)
    tot = np.zeros(3)
    for l in range(L):
        tot += orc.energy_binary(np.zeros(sh), Ct[..., l], a[..., l], u[..., l], sh)
    np.testing.assert_allclose(e, tot, rtol=1e-14)
