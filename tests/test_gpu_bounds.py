"""GPU primal-dual bounds of the binary model (asets_binary_bounds_dev) against the FP64
CPU restatement (oracle/bounds_oracle.py), which tests/test_gap_bounds.py pins by weak
duality, the tightness of the lambda descent and the reference fixed point's gap."""
import numpy as np
import pytest
import torch

import bounds_oracle as BO
import oracle as O
from asetsmatlabmaxflow_amd import solver, workloads

pytestmark = pytest.mark.gpu

REL = 1e-11   # FP64 on both sides; only the order of the volume sums differs


def _dev(a, cuda):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))).to(cuda)


def _ref(Cs, Ct, al, u, faces, shape, sweeps):
    g = lambda a: BO.grid(np.asarray(a, np.float32), shape)
    f = [g(x) for x in faces]
    P, tv, _, _ = BO.primal_bound(g(Cs), g(Ct), g(al), g(u), sweeps=sweeps)
    return P, tv, BO.dual(g(Cs), g(Ct), g(al), f)


def _check(cuda, Cs, Ct, al, u, faces, shape, sweeps):
    """faces: [bx, by, bz] (3D) or [bx, by] (2D, bz passed as NULL)."""
    d = [_dev(x, cuda) for x in (Cs, Ct, al, u, *faces)]
    if len(shape) == 2:
        d.append(None)
    got = solver.binary_bounds(*d, shape, sweeps=sweeps)
    shp3 = shape if len(shape) == 3 else (*shape, 1)
    f3 = list(faces) if len(shape) == 3 else [*faces, np.zeros(int(np.prod(shape)))]
    P, tv, D = _ref(Cs, Ct, al, u, f3, shp3, sweeps)
    for key, want in (("primal", P), ("primal_tv", tv), ("dual", D)):
        assert abs(got[key] - want) <= REL * abs(want), (key, got[key], want)
    assert got["dual"] <= got["primal"] <= got["primal_tv"] * (1 + 1e-12)
    return got


@pytest.mark.parametrize("its", [30, 400])
def test_bounds_of_solver_states(cuda, its):
    """States of the FP32 solve (oracle) on 24x20x16 ellipsoids after 30 and 400
    iterations."""
    shape = (24, 20, 16)
    print(f"bounds {shape} its={its}", flush=True)
    Cs, Ct, al, _ = workloads.binary3d_ellipsoids(shape, seed=5, count=3)
    r = O.binary(Cs, Ct, al, shape, its, 0.0, 0.35, 0.11, want_state=True)
    got = _check(cuda, Cs, Ct, al, r.u, [r.state[k] for k in ("bx", "by", "bz")], shape, 30)
    print(got)


@pytest.mark.parametrize("shape", [(17, 13, 9), (40, 30)])
def test_bounds_of_random_states(cuda, shape):
    """Random u in [-0.2, 1.2] (clipped inside), random faces, random alpha; 3D and 2D,
    with 0, 1 and 7 sweeps (an odd count ends on the second colour)."""
    print(f"bounds random {shape}", flush=True)
    rng = np.random.default_rng(11)
    n = int(np.prod(shape))
    Cs, Ct = rng.uniform(0, 3, n), rng.uniform(0, 3, n)
    al = rng.uniform(0.1, 1.0, n)
    u = rng.uniform(-0.2, 1.2, n)
    faces = [rng.normal(0, 0.5, n) for _ in range(len(shape))]
    for sweeps in (0, 1, 7):
        got = _check(cuda, Cs, Ct, al, u, faces, shape, sweeps)
        if sweeps == 0:
            assert got["primal"] == got["primal_tv"]


def test_bounds_full_size_bracket(cuda):
    """256^3 workload after 200 fixed iterations of the GPU solver: the bracket holds
    (dual <= primal <= voxel-centred primal), the lambda descent tightens the primal
    bound, and the relative width stays of the order the small volumes show."""
    shape = (256, 256, 256)
    print("bounds 256^3", flush=True)
    Cs, Ct, al, _ = workloads.binary3d_ellipsoids(shape, seed=1002)
    d = [solver.to_device(x) for x in (Cs, Ct, al)]
    r = solver.binary(*d, shape, 200, 0.0, workloads.BINARY3D_CC, workloads.BINARY3D_STEPS,
                      want_state=True)
    got = solver.binary_bounds(*d, r.u, r.state["bx"], r.state["by"], r.state["bz"], shape,
                               sweeps=30)
    print(got)
    assert got["dual"] <= got["primal"] < got["primal_tv"]
    assert 0.0 < got["width"] < 0.05
