"""Resume from a full iteration state (asets_binary_dev_resume): k iterations, then m
more from the returned state, equal k+m iterations in one solve bit for bit -- for the
fused kernel (fp32 and fp16 flow storage), the per-pass kernels, 2D and 3D; and resuming
from the CPU oracle's own state reproduces the oracle's longer run."""
import numpy as np
import pytest
import torch

import oracle as orc
from asetsmatlabmaxflow_amd import solver, workloads

pytestmark = pytest.mark.gpu


def _inputs(shape):
    if len(shape) == 3:
        Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=1002, count=4)
        return Cs, Ct, a, 0.35, 0.11
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=5)
    return Cs, Ct, a, 0.3, 0.16


@pytest.mark.parametrize("shape", [(45, 38, 29), (70, 53)])
@pytest.mark.parametrize("kernel,flow", [("fused", "f32"), ("fused", "f16"), ("passes", "f32")])
def test_resume_equals_uninterrupted(cuda, shape, kernel, flow):
    Cs, Ct, a, cc, st = _inputs(shape)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    k, m = 7, 9
    full = solver.binary(*d, shape, k + m, 0.0, cc, st, kernel=kernel, flow=flow,
                         want_state=True)
    part = solver.binary(*d, shape, k, 0.0, cc, st, kernel=kernel, flow=flow, want_state=True)
    state = dict(part.state, u=part.u)
    if len(shape) == 2:
        state.pop("bz")
    res = solver.binary_resume(*d, shape, state, m, 0.0, cc, st, kernel=kernel, flow=flow,
                               want_state=True)
    assert torch.equal(res.u.view(torch.int32), full.u.view(torch.int32))
    for f in ("ps", "pt", "bx", "by"):
        assert torch.equal(res.state[f].view(torch.int32), full.state[f].view(torch.int32)), f
    torch.testing.assert_close(res.cvg, full.cvg[k:], rtol=0, atol=0)


def test_resume_from_oracle_state_matches_oracle(cuda):
    shape = (33, 27, 19)
    Cs, Ct, a, cc, st = _inputs(shape)
    k, m = 5, 6
    part = orc.binary(Cs, Ct, a, shape, k, 0.0, cc, st, want_state=True)
    ref = orc.binary(Cs, Ct, a, shape, k + m, 0.0, cc, st)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    state = {f: torch.from_numpy(np.ascontiguousarray(v)).to(cuda)
             for f, v in dict(part.state, u=part.u).items()}
    res = solver.binary_resume(*d, shape, state, m, 0.0, cc, st, kernel="fused")
    assert np.array_equal(res.u.cpu().numpy().view(np.uint32), ref.u.view(np.uint32))


def test_resume_ignores_boundary_faces(cuda):
    """The index-0 faces are the zero-flux boundary: garbage there must not change u."""
    shape = (20, 18, 12)
    Cs, Ct, a, cc, st = _inputs(shape)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    part = solver.binary(*d, shape, 4, 0.0, cc, st, want_state=True)
    state = dict(part.state, u=part.u)
    ref = solver.binary_resume(*d, shape, state, 5, 0.0, cc, st)
    dirty = {k: v.clone() for k, v in state.items()}
    ny, nx, nz = shape
    dirty["by"].view(nz, nx, ny)[:, :, 0] = 7.0
    dirty["bx"].view(nz, nx, ny)[:, 0, :] = -3.0
    dirty["bz"].view(nz, nx, ny)[0, :, :] = 5.0
    got = solver.binary_resume(*d, shape, dirty, 5, 0.0, cc, st)
    assert torch.equal(got.u.view(torch.int32), ref.u.view(torch.int32))
