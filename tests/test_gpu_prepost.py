"""GPU data costs, weights and labels (include/asets.h, pre- and post-processing) against
the numpy oracle: labels exactly, costs and weights within one fp32 ulp (FP64 log/exp
implementations may differ in the last double bit)."""
import numpy as np
import pytest
import torch

import prepost_oracle as PO
from asetsmatlabmaxflow_amd import _native as nat
from asetsmatlabmaxflow_amd import prepost, solver, workloads

pytestmark = pytest.mark.gpu


def _ulps(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    ia, ib = a.view(np.int32).astype(np.int64), b.view(np.int32).astype(np.int64)
    return np.abs(ia - ib).max()


def test_loglik_costs(cuda):
    rng = np.random.default_rng(0)
    shape = (64, 48, 20)
    img = (rng.normal(0.5, 0.2, shape) ** 2).astype(np.float32)
    scr = rng.integers(0, 4, shape).astype(np.int32)
    scr[rng.random(shape) < 0.9] = 0                           # sparse scribbles
    ref = PO.cost_loglik(img, scr, 3)
    got = prepost.cost_loglik(solver.to_device(img),
                              torch.from_numpy(np.ascontiguousarray(scr.ravel(order="F"))).to(cuda),
                              3)
    assert _ulps(got.cpu().numpy(), ref) <= 1


def test_loglik_errors(cuda):
    img = solver.to_device(np.full((4, 4), 2.0, np.float32))
    scr = torch.ones(16, dtype=torch.int32, device=cuda)
    with pytest.raises(nat.AsetsError, match="constant"):
        prepost.cost_loglik(img, scr, 1)
    img = solver.to_device(np.arange(16, dtype=np.float32).reshape(4, 4))
    with pytest.raises(nat.AsetsError, match="label 2"):
        prepost.cost_loglik(img, scr, 2)


def test_neglog(cuda):
    p = np.random.default_rng(1).random(10000).astype(np.float32)
    p[:3] = [0.0, 1.0, 1e-30]
    got = prepost.cost_neglog(torch.from_numpy(p).to(cuda))
    assert _ulps(got.cpu().numpy(), PO.cost_neglog(p)) <= 1


@pytest.mark.parametrize("shape", [(40, 33, 9), (70, 51), (5, 1, 3)])
@pytest.mark.parametrize("power,normalize", [(1, True), (2, False)])
def test_alpha_contrast(cuda, shape, power, normalize):
    img = np.random.default_rng(2).normal(0.5, 0.2, shape).astype(np.float32)
    got = prepost.alpha_contrast(solver.to_device(img), shape, 0.1, 0.7, 15.0, power, normalize)
    ref = PO.alpha_contrast(img, 0.1, 0.7, 15.0, power, normalize).ravel(order="F")
    assert _ulps(got.cpu().numpy(), ref) <= 1


def test_alpha_reproduces_the_benchmark_generator(cuda):
    """The 256^3 workload's alpha (0.2 + 0.8 exp(-20 |grad I|^2), central differences) is
    the GPU weight with power 2 on the same image."""
    shape = (48, 40, 24)
    Cs, Ct, a, _, I = workloads.binary3d_ellipsoids(shape, seed=1002, count=4,
                                                    return_image=True)
    got = prepost.alpha_contrast(solver.to_device(I), shape, 0.2, 0.8, 20.0, 2, False)
    # the generator works in fp32 (numpy), the kernel in FP64
    np.testing.assert_allclose(got.cpu().numpy(), np.asarray(a).ravel(order="F"), rtol=5e-6,
                               atol=1e-7)


def test_labels(cuda):
    rng = np.random.default_rng(3)
    n, L = 5000, 5
    u = rng.random(n * L).astype(np.float32)
    u[:n][::7] = 0.5                                           # exact threshold values
    u[n:2 * n][::11] = u[:n][::11]                             # ties with the first label
    du = torch.from_numpy(u).to(cuda)
    for model, nl in ((0, 1), (1, L), (2, L)):
        got = prepost.labels(model, du, n, nl)
        np.testing.assert_array_equal(got.cpu().numpy(), PO.labels(model, u, n, nl))
