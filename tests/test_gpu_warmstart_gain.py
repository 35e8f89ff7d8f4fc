"""What the trained warm-start networks gain over the method without them, on volumes of
the training generator that no training run saw (seeds 5000+): the committed weights,
inference on the GPU, then the plain solver to the reference's stop rule (errBound 1e-4).

The space-to-depth network predicts the full state (u and the face flows; ps, pt paired
to them), which asets_binary_dev_resume continues: it must stop in fewer iterations than
the cold start with at least the cold start's label agreement with a converged solve and
an energy within 1e-3 (relative) of the converged one.  The full-resolution (u0, pt0)
network restarts the flows at 0, so it cannot stop sooner; its test checks only that its
start is no worse than the cold one."""
from pathlib import Path

import pytest
import torch

from asetsmatlabmaxflow_amd import solver, workloads
from asetsmatlabmaxflow_amd import warmstart as W

pytestmark = pytest.mark.gpu

WDIR = Path(W.__file__).resolve().parent / "weights"
CC, ST = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS


def _volume(seed, n=128):
    shape = (n, n, n)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=seed)
    return shape, [solver.to_device(x) for x in (Cs, Ct, a)]


def _quality(d, shape, u, ref_u, e_star):
    e = solver.energy_binary(*d, u, shape)[0]
    return (e - e_star) / abs(e_star), float(((u > 0.5) == (ref_u > 0.5)).float().mean())


@pytest.mark.parametrize("seed", [5000, 5001])
def test_full_state_warm_start_saves_iterations(cuda, seed):
    shape, d = _volume(seed)
    net = W.WarmStartNetS2D(cuda)
    net.load(WDIR / "warmstart_s2d.npz")
    ref = solver.binary(*d, shape, 2000, 0.0, CC, ST)
    e_star = solver.energy_binary(*d, ref.u, shape)[0]
    cold = solver.binary(*d, shape, 20000, 1e-4, CC, ST)
    st = net.predict_state(d[0], d[1], d[2], shape, 1)
    warm = solver.binary_resume(*d, shape, st, 20000, 1e-4, CC, ST)
    gap_c, lab_c = _quality(d, shape, cold.u, ref.u, e_star)
    gap_w, lab_w = _quality(d, shape, warm.u, ref.u, e_star)
    print(f"seed {seed}: cold {cold.iterations} it (gap {gap_c:.2e}, labels {lab_c:.6f}); "
          f"warm {warm.iterations} it (gap {gap_w:.2e}, labels {lab_w:.6f})")
    assert warm.iterations <= cold.iterations - 2
    assert lab_w >= lab_c
    assert gap_w < 1e-3


def test_u_pt_warm_start_no_worse_than_cold(cuda):
    shape, d = _volume(5002)
    net = W.WarmStartNet(cuda)
    net.load(WDIR / "warmstart_unet.npz")
    cold = solver.binary(*d, shape, 20000, 1e-4, CC, ST)
    u0, pt0 = net.forward(d[0], d[1], d[2], shape, 1)
    warm = solver.binary(*d, shape, 20000, 1e-4, CC, ST, u0=u0, pt0=pt0)
    assert warm.iterations <= cold.iterations
