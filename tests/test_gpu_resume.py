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


def _pair_reference(state, Cs, Ct, shape):
    """ps and pt paired to u and the faces in float32 torch operations, in the order
    asets_binary_dev_resume documents: dv = div b (faces at axis index 0 taken as 0),
    pt = Ct where u >= 1/2 else min(Cs - dv, Ct), ps = pt + dv, u clamped to [0, 1]."""
    ny, nx = shape[0], shape[1]
    nz = shape[2] if len(shape) == 3 else 1
    vol = (nz, nx, ny)
    z = torch.zeros((), dtype=torch.float32, device=state["u"].device)

    def low(f, dim):
        t = f.view(vol).clone()
        t.narrow(dim, 0, 1).zero_()
        return t

    def high(f, dim):
        t = torch.zeros_like(f.view(vol))
        n = vol[dim]
        t.narrow(dim, 0, n - 1).copy_(f.view(vol).narrow(dim, 1, n - 1))
        return t

    byl, bxl = low(state["by"], 2), low(state["bx"], 1)
    byh, bxh = high(state["by"], 2), high(state["bx"], 1)
    dv = ((byh - byl) + bxh) - bxl
    if len(shape) == 3:
        dv = (dv + high(state["bz"], 0)) - low(state["bz"], 0)
    dv = dv.reshape(-1)
    u = state["u"]
    u = torch.where(u < 0, z, torch.where(u > 1, z + 1, u))
    t = Cs - dv
    t = torch.where(t < Ct, t, Ct)
    pt = torch.where(u >= 0.5, Ct, t)
    return u, pt + dv, pt


@pytest.mark.parametrize("shape", [(45, 38, 29), (64, 40, 24), (70, 53)])
@pytest.mark.parametrize("kernel,flow", [("fused", "f32"), ("fused", "f16"), ("passes", "f32")])
def test_resume_pairs_ps_pt(cuda, shape, kernel, flow):
    """A resume given u and the faces only pairs ps and pt to them: the same bits as a
    resume from that state with ps and pt formed outside (torch, the documented order;
    in 3D also the network library's asets_nn_warm_state), u stretched past [0, 1] so
    the clamp is exercised."""
    from asetsmatlabmaxflow_amd.warmstart import pair_state

    Cs, Ct, a, cc, st = _inputs(shape)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    part = solver.binary(*d, shape, 6, 0.0, cc, st, kernel=kernel, flow=flow, want_state=True)
    faces = ("bx", "by", "bz") if len(shape) == 3 else ("bx", "by")
    pred = {"u": part.u * 1.25 - 0.1, **{f: part.state[f] for f in faces}}
    u_ref, ps_ref, pt_ref = _pair_reference(pred, d[0], d[1], shape)
    explicit = {"u": u_ref, "ps": ps_ref, "pt": pt_ref, **{f: pred[f] for f in faces}}
    if len(shape) == 3:
        nn = {k: v.clone() for k, v in pred.items()}
        pair_state(nn, d[0], d[1], shape)
        for f in ("u", "ps", "pt"):
            assert torch.equal(nn[f].view(torch.int32), explicit[f].view(torch.int32)), f
    kw = dict(kernel=kernel, flow=flow, want_state=True)
    r_exp = solver.binary_resume(*d, shape, explicit, 8, 0.0, cc, st, **kw)
    r_pair = solver.binary_resume(*d, shape, {k: v.clone() for k, v in pred.items()}, 8, 0.0,
                                  cc, st, **kw)
    assert torch.equal(r_pair.u.view(torch.int32), r_exp.u.view(torch.int32))
    for f in ("ps", "pt") + faces:
        assert torch.equal(r_pair.state[f].view(torch.int32), r_exp.state[f].view(torch.int32)), f
    torch.testing.assert_close(r_pair.cvg, r_exp.cvg, rtol=0, atol=0)
    # to an errBound: the same stop, and the stop's launch writes u (fused, ny % 4 == 0)
    e_exp = solver.binary_resume(*d, shape, explicit, 200, 1e-3, cc, st, kernel=kernel,
                                 flow=flow)
    e_pair = solver.binary_resume(*d, shape, dict(pred), 200, 1e-3, cc, st, kernel=kernel,
                                  flow=flow)
    assert e_pair.it_num == e_exp.it_num
    assert torch.equal(e_pair.u.view(torch.int32), e_exp.u.view(torch.int32))


def test_resume_pairing_needs_both_or_neither(cuda):
    shape = (20, 12, 8)
    Cs, Ct, a, cc, st = _inputs(shape)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    part = solver.binary(*d, shape, 3, 0.0, cc, st, want_state=True)
    state = dict(part.state, u=part.u)
    state.pop("ps")
    with pytest.raises(RuntimeError, match="both, or neither"):
        solver.binary_resume(*d, shape, state, 2, 0.0, cc, st)
