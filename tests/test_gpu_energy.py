"""GPU energy evaluator (asets_binary_energy_dev / asets_potts_energy_dev) against the
FP64 oracle: same definition, different summation order, so 1e-12 relative."""
import numpy as np
import pytest
import torch

import oracle as orc
from asetsmatlabmaxflow_amd import solver, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(37, 29, 11), (64, 64, 64), (130, 77), (1, 1, 1), (5, 1, 3)])
def test_binary_energy_matches_oracle(cuda, shape):
    rng = np.random.default_rng(len(shape) * 100 + shape[0])
    Cs, Ct = rng.uniform(0, 3, shape), rng.uniform(0, 3, shape)
    a, u = rng.uniform(0.1, 1, shape), rng.uniform(-0.2, 1.2, shape)
    Cs, Ct, a, u = (x.astype(np.float32) for x in (Cs, Ct, a, u))
    got = solver.energy_binary(*(solver.to_device(x) for x in (Cs, Ct, a, u)), shape)
    ref = orc.energy_binary(Cs, Ct, a, u, shape)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("shape,L", [((33, 20, 9), 4), ((40, 31), 3)])
def test_potts_energy_matches_oracle(cuda, shape, L):
    rng = np.random.default_rng(L)
    Ct, a, u = (rng.uniform(0, 1, shape + (L,)).astype(np.float32) for _ in range(3))
    got = solver.energy_potts(*(solver.to_device(x) for x in (Ct, a, u)), shape, L)
    ref = orc.energy_potts(Ct, a, u, shape, L)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)


def test_energy_of_solver_output_decreases_toward_the_minimum(cuda):
    """E(u_k) of the solver iterates: after many iterations the relaxed energy is below the
    energy of the initial cut, and the thresholded labelling's energy is close to it."""
    shape = (48, 40, 32)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=1002, count=4)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    u_init = solver.to_device((Cs >= Ct).astype(np.float32))
    e_init = solver.energy_binary(*d, u_init, shape)[0]
    r = solver.binary(*d, shape, 400, 0.0, 0.35, 0.11)
    e_end = solver.energy_binary(*d, r.u, shape)[0]
    e_thr = solver.energy_binary(*d, (r.u > 0.5).float(), shape)[0]
    assert e_end < e_init
    assert abs(e_thr - e_end) < 0.02 * abs(e_end)
