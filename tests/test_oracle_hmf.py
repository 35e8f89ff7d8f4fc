"""Pins of the hierarchical max-flow (full-flow) oracle, oracle.hmf, on the CPU.

The reference (maxflow/asetsHMF2D.m, asetsHMF3D.m) is MATLAB and has no test or
golden vector of its own (tests/runTestHMFLinearModel.m needs an image that is not
in the repository).  The oracle is pinned by:
  * the state after the first iteration from the zero state, derived by hand;
  * what the mathematics fixes at convergence: a two-leaf tree is the binary model
    and a flat tree the Potts model (both oracles pinned on their own), with the
    capacity of the full Euclidean norm (sqrt(2) times that of the C-MEX solvers'
    0.5-weighted norm); a branch with alpha = 0 changes nothing; a branch's label is
    the sum of its children's, the root's children sum to 1;
  * spatially constant costs, which keep every flow at exactly 0;
  * the shape constraint's documented special cases: a zero or short direction is
    no constraint, and a direction along an axis exempts the flows along it;
  * an independent restatement in the reference's own array form (faces stored as
    the reference's (D1-1) x D2 x D3 px arrays and so on, whole-array statements),
    bit for bit in fp32.
"""
import numpy as np
import pytest

import oracle as O

F32 = np.float32


# ---------------------------------------------------------------------------
# whole-array restatement (asetsHMF3D.m:191-336, asetsHMF2D.m:206-338)
# ---------------------------------------------------------------------------
def _sl(nd, k, lo, hi):
    idx = [slice(None)] * nd
    idx[k] = slice(lo, hi)
    return tuple(idx)


def hmf_arrays(parent, Ct, alpha, shape, num_it, steps, cc, l=None):
    D = tuple(shape)
    nd = len(D)
    M = len(parent)
    # MATLAB's steps and cc are doubles; with single arrays they act as single, except
    # the root's 1/cc .* ones(h.D), a double array rounded where it meets single data
    root_cc = F32(1.0 / float(cc))
    steps, cc = F32(steps), F32(cc)
    kids = [[c for c in range(M) if parent[c] == n] for n in range(M)]
    root = list(parent).index(-1)
    u = [np.zeros(D, F32) for _ in range(M)]
    pt = [np.zeros(D, F32) for _ in range(M)]
    div = [np.zeros(D, F32) for _ in range(M)]
    p = [[np.zeros(tuple(D[j] - (j == k) for j in range(nd)), F32) for k in range(nd)]
         for _ in range(M)]
    al = [np.broadcast_to(np.asarray(alpha[n], F32), D) for n in range(M)]
    ct = [None if Ct[n] is None else np.asarray(Ct[n], F32) for n in range(M)]
    L = [None] * M
    for n in range(M):
        if l is None or l[n] is None:
            continue
        q = [np.broadcast_to(np.asarray(x, F32), D) for x in l[n]]
        s2 = q[0] ** 2 + q[1] ** 2
        if nd == 3:
            s2 = s2 + q[2] ** 2
        den = np.sqrt(s2)
        m = den > F32(0.001)
        L[n] = [np.divide(x, den, out=np.zeros(D, F32), where=m) for x in q]

    def sink(n):
        for c in kids[n]:
            sink(c)
        if not kids[n]:
            pt[n] = np.minimum(ct[n], (pt[parent[n]] - div[n]) + u[n] / cc)
            return
        acc = np.full(D, root_cc) if n == root else (pt[parent[n]] - div[n]) + u[n] / cc
        for c in kids[n]:
            acc = ((acc + pt[c]) + div[c]) - u[c] / cc
        pt[n] = acc / F32(len(kids[n]) if n == root else 1 + len(kids[n]))

    for _ in range(num_it):
        sink(root)
        for n in range(M):
            if n != root:
                u[n] = u[n] + cc * ((pt[parent[n]] - div[n]) - pt[n])
        for n in range(M):
            if n == root:
                continue
            P = p[n]
            g = steps * (((div[n] + pt[n]) - pt[parent[n]]) - u[n] / cc)
            for k in range(nd):
                hi, lo = g[_sl(nd, k, 1, D[k])], g[_sl(nd, k, 0, D[k] - 1)]
                P[k] = (P[k] + hi) - lo if nd == 3 else P[k] + (hi - lo)
            ex = None
            if L[n] is not None:
                a = np.zeros(D, F32)
                a[_sl(nd, 0, 1, D[0])] = np.maximum(
                    np.where(P[0] < 0, P[0] * L[n][0][_sl(nd, 0, 1, D[0])], F32(0)), F32(0))
                a[_sl(nd, 0, 0, 1)] = 0
                a[_sl(nd, 0, 0, D[0] - 1)] += np.maximum(
                    np.where(P[0] > 0, P[0] * L[n][0][_sl(nd, 0, 0, D[0] - 1)], F32(0)), F32(0))
                for k in range(1, nd):
                    a[_sl(nd, k, 0, D[k] - 1)] += np.maximum(
                        np.where(P[k] > 0, P[k] * L[n][k][_sl(nd, k, 0, D[k] - 1)], F32(0)),
                        F32(0))
                    a[_sl(nd, k, 1, D[k])] += np.maximum(
                        np.where(P[k] < 0, P[k] * L[n][k][_sl(nd, k, 1, D[k])], F32(0)), F32(0))
                ex = []
                for k in range(nd):
                    llo, lhi = L[n][k][_sl(nd, k, 0, D[k] - 1)], L[n][k][_sl(nd, k, 1, D[k])]
                    alo, ahi = a[_sl(nd, k, 0, D[k] - 1)], a[_sl(nd, k, 1, D[k])]
                    e = np.where((P[k] > 0) & (llo > 0), llo * alo,
                                 np.where((P[k] < 0) & (lhi < 0), lhi * ahi, F32(0)))
                    ex.append(e.astype(F32))
                    P[k] = P[k] - ex[k]
            g = np.zeros(D, F32)
            if nd == 3:
                g[_sl(nd, 0, 1, D[0])] = P[0] * P[0]
                g[_sl(nd, 0, 0, 1)] = 0
                g[_sl(nd, 0, 0, D[0] - 1)] += P[0] * P[0]
            else:
                g[_sl(nd, 0, 0, D[0] - 1)] = P[0] * P[0]
                g[_sl(nd, 0, D[0] - 1, D[0])] = 0
                g[_sl(nd, 0, 1, D[0])] += P[0] * P[0]
            for k in range(1, nd):
                g[_sl(nd, k, 0, D[k] - 1)] += P[k] * P[k]
                g[_sl(nd, k, 1, D[k])] += P[k] * P[k]
            g = np.sqrt(g)
            keep = g <= al[n]
            g = np.where(keep, F32(1), np.divide(al[n], g, out=np.ones(D, F32), where=~keep))
            for k in range(nd):
                r = (F32(0.5) * P[k]) * (g[_sl(nd, k, 1, D[k])] + g[_sl(nd, k, 0, D[k] - 1)])
                P[k] = ex[k] + r if ex is not None else r
            d = np.zeros(D, F32)
            d[_sl(nd, 0, 0, D[0] - 1)] = P[0]
            d[_sl(nd, 0, D[0] - 1, D[0])] = 0
            d[_sl(nd, 0, 1, D[0])] -= P[0]
            for k in range(1, nd):
                d[_sl(nd, k, 0, D[k] - 1)] += P[k]
                d[_sl(nd, k, 1, D[k])] -= P[k]
            div[n] = d
    return u, pt, div, p, L


def _image(shape, nlab, seed, noise=0.4):
    """Blobs labelled 0..nlab-1 plus Gaussian noise; costs |img - k|."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    lab = np.zeros(shape, int)
    for k in range(1, nlab):
        c = [rng.uniform(0.25, 0.75) * s for s in shape]
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        lab[r2 < (0.22 * min(shape)) ** 2] = k
    img = lab + noise * rng.standard_normal(shape)
    return [np.abs(img - k) for k in range(nlab)]


# ---------------------------------------------------------------------------
# first iteration from the zero state, derived by hand
# ---------------------------------------------------------------------------
def test_first_iteration_closed_form():
    # root 0 -> [branch 1 -> (leaves 2, 3), leaf 4]: m = 2 children of the root
    rng = np.random.default_rng(1)
    shape = (9, 7)
    Ct = [None, None, rng.uniform(-1, 1, shape), rng.uniform(-1, 1, shape),
          rng.uniform(0.1, 1, shape)]
    cc, steps = 0.5, 0.1
    r = O.hmf([-1, 0, 1, 1, 0], Ct, [0, 10.0, 10.0, 10.0, 10.0], shape, 1, steps, cc,
              precision="f64")
    # sink flows, bottom-up: leaves min(Ct, P.pt_old - 0 + 0) = min(Ct, 0); the
    # branch (0 + sum of its children's new pt) / 3; the root (1/cc + sum of its
    # children's new pt) / 2; labels u = cc*(P.pt - pt) since div = u = 0 before
    pl = {n: np.minimum(Ct[n], 0) for n in (2, 3, 4)}
    pB = (pl[2] + pl[3]) / 3
    pR = (1.0 / cc + pB + pl[4]) / 2
    np.testing.assert_array_equal(r.pt[2], pl[2])
    np.testing.assert_array_equal(r.pt[4], 0 * Ct[4])
    np.testing.assert_allclose(r.pt[1], pB, atol=1e-15)
    np.testing.assert_allclose(r.pt[0], pR, rtol=1e-15)
    np.testing.assert_allclose(r.u[1], cc * (pR - pB), rtol=1e-14)
    np.testing.assert_allclose(r.u[4], cc * pR, rtol=1e-14)
    np.testing.assert_allclose(r.u[2], cc * (pB - pl[2]), atol=1e-15)
    np.testing.assert_allclose(r.u[3], cc * (pB - pl[3]), atol=1e-15)
    # spatial flows of leaf 2 (alpha large: no projection): faces are the differences
    # of g = steps*(pt - P.pt - u/cc) across them, div its discrete Laplacian
    g = steps * ((r.pt[2] - r.pt[1]) - r.u[2] / cc)
    f0 = np.zeros(shape)
    f0[1:, :] = g[1:, :] - g[:-1, :]
    f1 = np.zeros(shape)
    f1[:, 1:] = g[:, 1:] - g[:, :-1]
    np.testing.assert_allclose(r.faces[2, 0], f0, atol=1e-15)
    np.testing.assert_allclose(r.faces[2, 1], f1, atol=1e-15)
    lap = np.zeros(shape)
    lap[:-1, :] += f0[1:, :]
    lap -= f0
    lap[:, :-1] += f1[:, 1:]
    lap -= f1
    np.testing.assert_allclose(r.div[2], lap, atol=1e-14)
    # costs: leaves sum u*Ct, non-root nodes sum u*div, totals over subtrees
    own_d = [0, 0, (r.u[2] * Ct[2]).sum(), (r.u[3] * Ct[3]).sum(), (r.u[4] * Ct[4]).sum()]
    own_r = [0] + [(r.u[n] * r.div[n]).sum() for n in range(1, 5)]
    np.testing.assert_allclose(r.data_cost[1], own_d[2] + own_d[3], rtol=1e-12)
    np.testing.assert_allclose(r.data_cost[0], sum(own_d), rtol=1e-12)
    np.testing.assert_allclose(r.regu_cost[0], sum(own_r), rtol=1e-10, atol=1e-14)


# ---------------------------------------------------------------------------
# convergence: what the model reduces to
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(40, 36), (18, 16, 14)])
def test_two_leaf_tree_is_the_binary_model(shape):
    Ct0, Ct1 = _image(shape, 2, seed=3, noise=0.35)
    ah = 0.25
    r = O.hmf([-1, 0, 0], [None, Ct1, Ct0], [0, ah, ah], shape, 600, 0.1, 0.3,
              precision="f64")
    b = O.binary(Ct0, Ct1, np.full(shape, np.sqrt(2) * ah), shape, 4000, 0.0, 0.3, 0.16,
                 precision="f64")
    ub = b.u.reshape(shape, order="F")
    d = np.abs(r.u[1] - ub)
    # capacity alpha per leaf in the Euclidean norm == sqrt(2)*alpha of the binary
    # model's 0.5-weighted norm; a 10 % error in the capacity doubles mean |du|
    assert d.mean() < 2.5e-4 and d.max() < 0.05, (d.mean(), d.max())
    np.testing.assert_allclose(r.u[1] + r.u[2], 1.0, atol=1e-12)


def test_flat_tree_is_the_potts_model_and_zero_alpha_branch_is_neutral():
    shape = (36, 32)
    Ct = _image(shape, 3, seed=5)
    ah = 0.3
    r = O.hmf([-1, 0, 0, 0], [None] + Ct, [0, ah, ah, ah], shape, 800, 0.1, 1 / 3,
              precision="f64")
    p = O.potts(np.stack(Ct, -1), np.full(shape + (3,), ah / np.sqrt(2)), shape, 3, 4000,
                0.0, 0.3, 0.11, precision="f64")
    up = p.u.reshape(shape + (3,), order="F")
    for k in range(3):
        d = np.abs(r.u[1 + k] - up[..., k])
        assert d.mean() < 4e-4 and d.max() < 0.05, (k, d.mean(), d.max())
    # root -> [B -> (b1, b2), c] with alpha_B = 0 minimises the same energy
    rn = O.hmf([-1, 0, 1, 1, 0], [None, None, Ct[0], Ct[1], Ct[2]], [0, 0.0, ah, ah, ah],
               shape, 800, 0.1, 1 / 3, precision="f64")
    for k, n in enumerate([2, 3, 4]):
        assert np.abs(rn.u[n] - r.u[1 + k]).max() < 1e-4
    np.testing.assert_allclose(rn.u[1], rn.u[2] + rn.u[3], atol=1e-4)
    np.testing.assert_allclose(rn.u[1] + rn.u[4], 1.0, atol=1e-12)


def test_nested_tree_invariants_with_branch_regularisation():
    shape = (14, 12, 10)
    Ct = _image(shape, 4, seed=9)
    # root -> [B1 -> (a, b), B2 -> (c, d)], every node regularised
    r = O.hmf([-1, 0, 0, 1, 1, 2, 2], [None, None, None] + Ct, [0, 0.1, 0.1, 0.2, 0.2, 0.2, 0.2],
              shape, 500, 0.1, 0.25, precision="f64")
    np.testing.assert_allclose(r.u[1] + r.u[2], 1.0, atol=1e-10)
    np.testing.assert_allclose(r.u[1], r.u[3] + r.u[4], atol=2e-3)
    np.testing.assert_allclose(r.u[2], r.u[5] + r.u[6], atol=2e-3)
    assert r.u.min() > -0.02 and r.u.max() < 1.02


def test_constant_costs_keep_flows_at_zero():
    shape = (10, 9, 8)
    costs = [0.7, 0.2, 0.5]
    Ct = [None] + [np.full(shape, c) for c in costs]
    r = O.hmf([-1, 0, 0, 0], Ct, [0, 0.2, 0.2, 0.2], shape, 300, 0.1, 0.3)
    assert not r.faces.any() and not r.div.any()
    # the fixed point is the cheapest label everywhere
    np.testing.assert_allclose(r.u[2], 1.0, atol=1e-3)
    np.testing.assert_allclose(r.u[1], 0.0, atol=1e-3)
    np.testing.assert_allclose(r.u[3], 0.0, atol=1e-3)
    np.testing.assert_allclose(r.data_cost[0], 0.2 * np.prod(shape), rtol=1e-3)
    assert r.regu_cost[0] == 0.0


# ---------------------------------------------------------------------------
# shape constraint (asetsHMF3D.m:214-261)
# ---------------------------------------------------------------------------
def _leaf_pair(shape, seed):
    Ct0, Ct1 = _image(shape, 2, seed=seed)
    return [None, Ct1, Ct0], [0, 0.2, 0.2]


@pytest.mark.parametrize("shape", [(16, 14), (10, 9, 8)])
def test_zero_or_short_direction_is_no_constraint(shape):
    Ct, al = _leaf_pair(shape, 11)
    nd = len(shape)
    free = O.hmf([-1, 0, 0], Ct, al, shape, 20, 0.1, 0.3)
    short = [1e-4] + [0.0] * (nd - 1)   # |l| <= 0.001 normalises to 0
    for l in ([0.0] * nd, short):
        r = O.hmf([-1, 0, 0], Ct, al, shape, 20, 0.1, 0.3, l=[None, l, l])
        np.testing.assert_array_equal(r.u, free.u)
        np.testing.assert_array_equal(r.faces, free.faces)
        assert not r.l[1].any()


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_axis_direction_exempts_the_flows_along_it(sign):
    # l = +-e1: the dim-1 flows pointing along l take no part in the projection, so
    # after one iteration they equal the stepped flows (the same run with a capacity
    # that never binds); a direction of length 3 normalises to the unit one
    shape = (16, 14)
    Ct, _ = _leaf_pair(shape, 13)
    # costs partly negative, so that the first iteration already moves flow
    Ct = [None, Ct[1] - 0.5, Ct[2] - 0.5]
    l = [sign * 3.0, 0.0]
    r = O.hmf([-1, 0, 0], Ct, [0, 0.05, 0.05], shape, 1, 0.5, 0.3, l=[None, l, l])
    unit = O.hmf([-1, 0, 0], Ct, [0, 0.05, 0.05], shape, 1, 0.5, 0.3,
                 l=[None, [sign, 0.0], [sign, 0.0]])
    np.testing.assert_array_equal(r.faces, unit.faces)
    np.testing.assert_array_equal(r.l[1][0], np.full(shape, sign, np.float32))
    stepped = O.hmf([-1, 0, 0], Ct, [0, 1e30, 1e30], shape, 1, 0.5, 0.3)
    f = r.faces[1, 0]
    along = (f * sign) > 0
    assert along.sum() > 20
    np.testing.assert_array_equal(f[along], stepped.faces[1, 0][along])
    # the other faces are projected: some of them are cut back
    assert np.abs(f[~along]).max() < np.abs(stepped.faces[1, 0][~along]).max()


# ---------------------------------------------------------------------------
# independent whole-array restatement, fp32, bit for bit
# ---------------------------------------------------------------------------
def test_root_sink_term_is_the_double_quotient_rounded():
    """asetsHMF3D.m:320 sets the root's sink flow to 1/cc .* ones(h.D) in double; it is
    rounded to single where it meets the children's single pt.  For cc = 0.12 that value,
    single(1/0.12) = 8.333333, is one ulp below the single quotient 1/single(0.12) =
    8.333334.  One iteration from zero with constant costs keeps every spatial flow and
    div at 0, so the root's pt is (single(1/cc) + sum of the leaves' pt) / 2, rounded in
    that order, and the leaves' pt = min(Ct, 0)."""
    cc = 0.12
    assert F32(1.0 / cc) != F32(1) / F32(cc)
    shape = (5, 4)
    c1, c2 = F32(-0.25), F32(-0.5)
    r = O.hmf([-1, 0, 0], [None, np.full(shape, c1), np.full(shape, c2)], [0, 1.0, 1.0],
              shape, 1, 0.1, cc)
    want = ((F32(1.0 / cc) + c1) + c2) / F32(2)
    np.testing.assert_array_equal(r.pt[1], np.full_like(r.pt[1], c1))
    np.testing.assert_array_equal(r.pt[0], np.full_like(r.pt[0], want))
    wrong = ((F32(1) / F32(cc) + c1) + c2) / F32(2)
    assert want != wrong


@pytest.mark.parametrize("shape,shape_constraint", [((13, 11), False), ((13, 11), True),
                                                    ((9, 8, 7), False), ((9, 8, 7), True)])
def test_array_restatement_bitwise(shape, shape_constraint):
    rng = np.random.default_rng(17)
    Ct = _image(shape, 3, seed=21)
    parent = [-1, 0, 1, 1, 0]
    Ctn = [None, None, Ct[0], Ct[1], Ct[2]]
    alpha = [0, 0.15, rng.uniform(0.05, 0.3, shape), 0.2, 0.25]
    l = None
    if shape_constraint:
        nd = len(shape)
        l = [None, [rng.standard_normal(shape) for _ in range(nd)], None,
             [rng.standard_normal(shape) * (rng.random(shape) > 0.2) for _ in range(nd)], None]
    its, steps, cc = 15, 0.2, 0.35
    r = O.hmf(parent, Ctn, alpha, shape, its, steps, cc, l=l)
    u, pt, div, p, L = hmf_arrays(parent, Ctn, alpha, shape, its, steps, cc, l=l)
    nd = len(shape)
    for n in range(len(parent)):
        np.testing.assert_array_equal(r.u[n], u[n])
        np.testing.assert_array_equal(r.pt[n], pt[n])
        np.testing.assert_array_equal(r.div[n], div[n])
        if n == 0:
            continue
        for k in range(nd):
            own = r.faces[n, k]
            np.testing.assert_array_equal(own[_sl(nd, k, 1, shape[k])], p[n][k])
            assert not own[_sl(nd, k, 0, 1)].any()
        if L[n] is not None:
            for k in range(nd):
                np.testing.assert_array_equal(r.l[n][k], L[n][k])
    assert np.abs(r.faces).max() > 0.01


def test_bad_trees_are_rejected():
    shape = (4, 4)
    c = np.ones(shape)
    for parent in ([-1, -1, 0], [1, 0, 0], [-1], [-1, 0, 5], [-1, 2, 1]):
        Ct = [None] + [c] * (len(parent) - 1)
        with pytest.raises(O.OracleError):
            O.hmf(parent, Ct, [0.1] * len(parent), shape, 1, 0.1, 0.3)
