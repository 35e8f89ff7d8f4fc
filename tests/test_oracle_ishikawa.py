"""Pins of the Ishikawa restatement in the oracle (orc_ishikawa_f32/f64), which follows
maxflow/asetsIshikawa3D_mex.c and asetsIshikawa2D_mex.c: the closed-form initial cut,
the one-voxel fixed point, spatially uniform costs (every voxel runs the one-voxel
recurrence, bit for bit), Nz = 1 equals the 2D solver, two labels equal the binary model
at convergence (the binary oracle is pinned on its own), and the symmetry of the
converged cut functions under reversal of the label order (which fixes the middle-layer
sink-flow formula)."""
import numpy as np
import pytest

import oracle as orc


def _random(shape, L, seed, smooth=True):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    Ct = rng.uniform(0, 1, shape + (L,))
    if smooth:
        Ct = gaussian_filter(Ct, (1.0,) * len(shape) + (0,))
    return Ct.astype(np.float32)


def test_initial_cut_closed_form():
    sh, L = (5, 4, 3), 4
    Ct = _random(sh, L, 0, smooth=False)
    Ct[0, 0, 0, :] = [0.5, 0.2, 0.2, 0.9]       # tie: the first arg-min wins
    a = np.full(sh + (L - 1,), 0.3, np.float32)
    r = orc.ishikawa(Ct, a, sh, L, 0, 0.0, 0.3, 0.11, want_state=True)
    N = int(np.prod(sh))
    ct = Ct.reshape(N, L, order="F")
    am = ct.argmin(1)
    u = r.u.reshape(L - 1, N)
    for l in range(L - 1):
        np.testing.assert_array_equal(u[l], (l < am).astype(np.float32))
    pt = r.state["pt"].reshape(L, N)
    for l in range(L - 1):
        np.testing.assert_array_equal(pt[l], ct.min(1))
    assert not pt[L - 1].any()   # the top layer is never initialised


@pytest.mark.parametrize("ct", [[3.0, 1.0, 2.0, 4.0], [0.2, 0.9, 0.8], [2.0, 1.5]])
def test_one_voxel_fixed_point(ct):
    """One voxel: no spatial flow; the converged cut functions are the cumulative
    indicator of the cheapest label and every sink flow is min Ct.  errBound -1: the
    error can be exactly 0 at an early iteration (and 0 <= 0 would stop the 3D solver)."""
    L = len(ct)
    Ct = np.array(ct, np.float64).reshape(1, 1, 1, L)
    a = np.full((1, 1, 1, L - 1), 0.5)
    r = orc.ishikawa(Ct, a, (1, 1, 1), L, 3000, -1.0, 0.3, 0.11, precision="f64",
                     want_state=True)
    m = int(np.argmin(ct))
    np.testing.assert_allclose(r.u, [1.0 if l < m else 0.0 for l in range(L - 1)], atol=1e-6)
    np.testing.assert_allclose(r.state["pt"], min(ct), atol=1e-6)


@pytest.mark.parametrize("is3d", [True, False])
def test_uniform_costs_follow_the_one_voxel_run(is3d):
    L = 4
    ct = np.array([0.7, 0.3, 0.5, 0.9], np.float32)
    sh = (6, 5, 4) if is3d else (6, 5)
    Ct = np.broadcast_to(ct, sh + (L,)).astype(np.float32)
    a = np.full(sh + (L - 1,), 0.25, np.float32)
    r = orc.ishikawa(Ct, a, sh, L, 40, 0.0, 0.3, 0.11)
    one = orc.ishikawa(ct.reshape((1,) * len(sh) + (L,)), a.reshape(-1)[:L - 1],
                       (1,) * len(sh), L, 40, 0.0, 0.3, 0.11)
    N = int(np.prod(sh))
    u = r.u.reshape(L - 1, N)
    for l in range(L - 1):
        assert np.all(u[l].view(np.uint32) == one.u[l:l + 1].view(np.uint32))


def test_nz1_equals_2d():
    sh, L = (9, 7), 3
    Ct = _random(sh, L, 3)
    a = np.full(sh + (L - 1,), 0.2, np.float32)
    r2 = orc.ishikawa(Ct, a, sh, L, 30, 0.0, 0.3, 0.11)
    r3 = orc.ishikawa(Ct.reshape(9, 7, 1, L), a.reshape(9, 7, 1, L - 1), (9, 7, 1), L, 30,
                      0.0, 0.3, 0.11)
    assert np.array_equal(r2.u.view(np.uint32), r3.u.view(np.uint32))


def test_stop_rules_2d_strict_3d_inclusive():
    """One voxel with Ct = (3, 1, 2, 4), cc = 1/4: the error is exactly 0 at iteration 1
    (every sink flow equal for one step).  The 3D solver stops there (0 <= 0), the 2D
    solver does not (0 < 0 is false) and runs on."""
    ct = np.array([3.0, 1.0, 2.0, 4.0], np.float32)
    L = 4
    for sh, want in (((1, 1, 1), 1), ((1, 1), 5)):
        r = orc.ishikawa(ct.reshape(sh + (L,)), np.full(sh + (L - 1,), 0.5, np.float32), sh,
                         L, 5, 0.0, 0.25, 0.11)
        assert r.cvg_ref[1] == 0.0
        assert r.it_num == want, sh


def test_two_labels_equal_binary_at_convergence():
    """L = 2 is the binary model with Cs = Ct_0 and Ct = Ct_1 (the expressions differ in
    association and the top sink flow starts at 0, so only the limits agree)."""
    sh = (12, 10, 6)
    Ct = _random(sh, 2, 4)
    a = np.full(sh + (1,), 0.15, np.float32)
    ri = orc.ishikawa(Ct, a, sh, 2, 3000, 0.0, 0.3, 0.11, precision="f64")
    rb = orc.binary(Ct[..., 0], Ct[..., 1], a[..., 0], sh, 3000, 0.0, 0.3, 0.11,
                    precision="f64")
    np.testing.assert_allclose(ri.u, rb.u, atol=2e-3)
    assert ((ri.u > 0.5) == (rb.u > 0.5)).mean() > 0.999


def _ordered_costs(shape, L, seed):
    """Piecewise-constant intensity classes with noise: Ct_l = (I - l/(L-1))^2 / 0.02,
    a problem with a clear, unique minimiser."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    cls = gaussian_filter(rng.uniform(0, L, shape), 1.5)
    cls = np.clip(np.round((cls - cls.min()) / (np.ptp(cls) + 1e-9) * (L - 1)), 0, L - 1)
    I = cls / (L - 1) + rng.normal(0, 0.08, shape)
    mu = np.arange(L) / (L - 1)
    return ((I[..., None] - mu) ** 2 / 0.02).astype(np.float32)


def test_label_reversal_symmetry_at_convergence():
    """Reversing the label order maps the cut functions u_l to 1 - u_{L-2-l} at the
    solution of the convex problem."""
    sh, L = (14, 12, 6), 4
    Ct = _ordered_costs(sh, L, 5)
    a = np.full(sh + (L - 1,), 0.3, np.float32)
    r = orc.ishikawa(Ct, a, sh, L, 3000, -1.0, 0.3, 0.11, precision="f64")
    rr = orc.ishikawa(Ct[..., ::-1], a[..., ::-1], sh, L, 3000, -1.0, 0.3, 0.11,
                      precision="f64")
    N = int(np.prod(sh))
    u, ur = r.u.reshape(L - 1, N), rr.u.reshape(L - 1, N)
    for l in range(L - 1):
        np.testing.assert_allclose(ur[l], 1.0 - u[L - 2 - l], atol=5e-3)
    # and the labelling is the expected one almost everywhere
    lab = (u > 0.5).sum(0)
    assert (lab == np.asarray(Ct).reshape(N, L, order="F").argmin(1)).mean() > 0.95


def test_fp32_tracks_fp64():
    sh, L = (11, 8, 5), 5
    Ct = _random(sh, L, 6)
    a = np.full(sh + (L - 1,), 0.2, np.float32)
    r32 = orc.ishikawa(Ct, a, sh, L, 100, 0.0, 0.3, 0.11)
    r64 = orc.ishikawa(Ct, a, sh, L, 100, 0.0, 0.3, 0.11, precision="f64")
    np.testing.assert_allclose(r32.u, r64.u, atol=1e-4)
    np.testing.assert_allclose(r32.cvg64, r64.cvg64, rtol=1e-3, atol=1e-7)
