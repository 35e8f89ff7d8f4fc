"""Ishikawa model on the GPU (label-parallel fused flow kernel + per-voxel sink-flow chain)
against the FP32 reference-order oracle: u bit for bit at a fixed iteration count, the
convergence trace against the oracle's FP64 sum, both stop rules (3D <=, 2D <), the mex
boundary and argument errors."""
import numpy as np
import pytest
import torch

import oracle as orc
from asetsmatlabmaxflow_amd import _native as nat
from asetsmatlabmaxflow_amd import mex, solver

pytestmark = pytest.mark.gpu


def _costs(shape, L, seed):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    Ct = gaussian_filter(rng.uniform(0, 1, shape + (L,)), (1.0,) * len(shape) + (0,))
    a = np.stack([np.full(shape, 0.3 / (l + 1)) for l in range(L - 1)], -1)
    return Ct.astype(np.float32), a.astype(np.float32)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32),
                          np.asarray(b, np.float32).view(np.uint32))


def _run(Ct, a, shape, L, iters, eb, cc, st, **kw):
    d = [solver.to_device(x) for x in (Ct, a)]
    return solver.ishikawa(*d, shape, L, iters, eb, cc, st, **kw)


@pytest.mark.parametrize("shape,L", [((45, 38, 29), 2), ((45, 38, 29), 3), ((33, 70, 17), 4),
                                     ((20, 21, 22), 5), ((37, 19, 11), 7), ((70, 53), 3),
                                     ((129, 40), 4), ((64, 64), 9)])
def test_fixed_iterations_bitwise(cuda, shape, L):
    Ct, a = _costs(shape, L, sum(shape) + L)
    iters = 25
    ref = orc.ishikawa(Ct, a, shape, L, iters, 0.0, 0.75, 0.11 if len(shape) == 3 else 0.16)
    r = _run(Ct, a, shape, L, iters, 0.0, 0.75, 0.11 if len(shape) == 3 else 0.16)
    assert r.it_num == ref.it_num == iters
    assert _bits_equal(r.u.cpu().numpy(), ref.u)
    np.testing.assert_allclose(r.cvg.cpu().numpy(), ref.cvg64, rtol=1e-6, atol=0)


@pytest.mark.parametrize("tx,zc", [(8, 8), (32, 4), (16, 3)])
def test_tilings_bitwise(cuda, tx, zc):
    shape, L = (70, 45, 23), 4
    Ct, a = _costs(shape, L, 11)
    ref = orc.ishikawa(Ct, a, shape, L, 12, 0.0, 0.75, 0.1)
    r = _run(Ct, a, shape, L, 12, 0.0, 0.75, 0.1, tile_x=tx, zchunk=zc)
    assert _bits_equal(r.u.cpu().numpy(), ref.u)


@pytest.mark.parametrize("shape", [(40, 33, 21), (90, 61)])
def test_stop_rule_matches_oracle(cuda, shape):
    L = 4
    Ct, a = _costs(shape, L, 3)
    full = orc.ishikawa(Ct, a, shape, L, 400, 0.0, 0.75, 0.1, stop_rule=1)
    eb = float(full.cvg64[60]) * 1.0001   # crossed around iteration 60
    ref = orc.ishikawa(Ct, a, shape, L, 400, eb, 0.75, 0.1, stop_rule=1)
    r = _run(Ct, a, shape, L, 400, eb, 0.75, 0.1)
    assert ref.it_num < 400
    assert r.it_num == ref.it_num
    assert _bits_equal(r.u.cpu().numpy(), ref.u)


def test_exact_zero_error_stop_rules(cuda):
    """One voxel, Ct = (3, 1, 2, 4), cc = 1/4: the error is exactly 0 at iteration 1; the
    3D solver stops there (0 <= 0), the 2D solver runs on (0 < 0 is false)."""
    ct = np.array([3.0, 1.0, 2.0, 4.0], np.float32)
    for sh, want in (((1, 1, 1), 1), ((1, 1), 5)):
        Ct, a = ct.reshape(sh + (4,)), np.full(sh + (3,), 0.5, np.float32)
        ref = orc.ishikawa(Ct, a, sh, 4, 5, 0.0, 0.25, 0.11)
        r = _run(Ct, a, sh, 4, 5, 0.0, 0.25, 0.11)
        assert ref.it_num == want and r.it_num == want, sh
        assert _bits_equal(r.u.cpu().numpy(), ref.u)


@pytest.mark.parametrize("shape,L", [((45, 38, 29), 3), ((33, 70, 17), 5), ((70, 53), 4)])
def test_deferred_stop_test_same_bits(cuda, shape, L):
    """errBound 0 runs each iteration's stop test in the first workgroup of the next flow
    launch (3D: cvg <= 0, 2D: cvg < 0, never); a negative errBound keeps the finalize
    kernel after every iteration.  u, cvg and the iteration count are the same bits."""
    Ct, a = _costs(shape, L, 5 * L + 1)
    r0 = _run(Ct, a, shape, L, 25, 0.0, 0.75, 0.1)
    rn = _run(Ct, a, shape, L, 25, -1.0, 0.75, 0.1)
    assert r0.it_num == rn.it_num == 25
    assert _bits_equal(r0.cvg.cpu().numpy(), rn.cvg.cpu().numpy())
    assert _bits_equal(r0.u.cpu().numpy(), rn.u.cpu().numpy())
    assert bool((r0.cvg > 0).all())


def test_mex_boundary(cuda):
    sh3, L = (31, 22, 9), 4
    Ct, a = _costs(sh3, L, 5)
    a_full = np.concatenate([a, a[..., -1:]], -1)   # 3D reference test passes L fields
    pars = [31, 22, 9, L, 30, 0.0, 0.75, 0.1]
    u, cvg, it, t = mex.asetsIshikawa3D_mex(Ct, a_full, pars)
    ref = orc.ishikawa(Ct, a, sh3, L, 30, 0.0, 0.75, 0.1)
    assert u.shape == sh3 + (L - 1,) and u.dtype == np.float32
    assert _bits_equal(u.ravel(order="F"), ref.u)
    assert it.dtype == np.uint16 and int(it.ravel()[0]) == 30
    sh2 = (40, 35)
    Ct2, a2 = _costs(sh2, 3, 6)
    u2, cvg2, it2, _ = mex.asetsIshikawa2D_mex(Ct2, a2, [40, 35, 3, 20, 0.0, 0.75, 0.16])
    ref2 = orc.ishikawa(Ct2, a2, sh2, 3, 20, 0.0, 0.75, 0.16)
    assert u2.shape == sh2 + (2,)
    assert _bits_equal(u2.ravel(order="F"), ref2.u)


def test_errors(cuda):
    sh, L = (8, 8, 4), 3
    Ct, a = _costs(sh, L, 1)
    a[3, 3, 1, 1] = 0.0
    with pytest.raises(nat.AsetsError, match="alpha"):
        _run(Ct, a, sh, L, 3, 0.0, 0.75, 0.1)
    with pytest.raises(nat.AsetsError, match="2 labels"):
        _run(Ct[..., :1], a[..., :1], sh, 1, 3, 0.0, 0.75, 0.1)
    with pytest.raises(mex.MexError, match="alpha"):
        mex.asetsIshikawa3D_mex(Ct, a[..., :1], [8, 8, 4, L, 3, 0.0, 0.75, 0.1])
