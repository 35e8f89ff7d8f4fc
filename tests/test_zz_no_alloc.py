"""asets_binary_dev_noalloc (the allocation-free binary solve) against asets_binary_dev,
both called through the C ABI: the same u, cvg and it_num bit for bit, on calls in a row
that share the library's scratch array (each smaller than the one before, so a call
that read what the previous one left would show), and the documented rejections."""
import ctypes as C

import numpy as np
import pytest
import torch

from asetsmatlabmaxflow_amd import _native as nat
from asetsmatlabmaxflow_amd import solver, workloads

pytestmark = pytest.mark.gpu


def _solve(entry, d, shape, max_iter, err_bound, cc, steps, *, flow="f32", kernel="auto",
           u0=None, pt0=None, poll_every=0):
    ndim = len(shape)
    ny, nx = shape[0], shape[1]
    nz = shape[2] if ndim == 3 else 1
    N = ny * nx * nz
    dev = d[0].device
    u = torch.full((N,), 7.0, dtype=torch.float32, device=dev)
    cvg = torch.full((max(max_iter, 1),), -1.0, dtype=torch.float32, device=dev)
    itn = torch.full((1,), -5, dtype=torch.int32, device=dev)
    secs = C.c_double(0.0)
    opts = nat.make_opts(torch.cuda.current_stream().cuda_stream, solver._FLOWS[flow],
                         solver._KERNELS[kernel], poll_every, 0, 0, None)
    p = nat.make_params(max_iter, err_bound, cc, steps)
    fn = getattr(nat.lib(), entry)
    nat.check(fn(ndim, *(t.data_ptr() for t in d), ny, nx, nz, C.byref(p),
                 None if u0 is None else u0.data_ptr(), None if pt0 is None else pt0.data_ptr(),
                 u.data_ptr(), cvg.data_ptr(), itn.data_ptr(), C.byref(secs), C.byref(opts)))
    torch.cuda.synchronize()
    return u, cvg[:max_iter], int(itn.item())


def _same(a, b):
    return torch.equal(a.view(torch.int32), b.view(torch.int32))


CASES = [
    # shape, max_iter, err_bound, flow, kernel, warm
    ((128, 128, 128), 40, 0.0, "f32", "fused", False),     # the largest fp32 fused solve
    ((96, 70, 40), 60, 2e-4, "f32", "fused", False),        # stops early (polls)
    ((80, 64, 33), 50, 0.0, "f16", "fused", False),
    ((70, 45, 20), 30, 0.0, "f32", "passes", False),
    ((64, 40, 16), 25, 0.0, "f32", "fused", True),          # warm start
    ((90, 77), 80, 0.0, "f32", "fused", False),             # 2D
    ((33, 21, 7), 20, 0.0, "f32", "fused", False),
]


def test_noalloc_matches_binary_dev_calls_in_a_row(cuda):
    cc, steps = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    for k, (shape, its, eb, flow, kernel, warm) in enumerate(CASES):
        print(f"noalloc case {k}: shape={shape} its={its} eb={eb} {flow} {kernel} "
              f"warm={warm}", flush=True)
        if len(shape) == 3:
            Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=300 + k, count=4)
        else:
            Cs, Ct, a = workloads.random_smooth_binary(shape, seed=300 + k, alpha=0.2)
        d = [solver.to_device(x, cuda) for x in (Cs, Ct, a)]
        u0 = pt0 = None
        if warm:
            w = solver.binary(*d, shape, 7, 0.0, cc, steps)
            u0 = w.u.clone()
            pt0 = torch.minimum(d[0], d[1])
        ref = _solve("asets_binary_dev", d, shape, its, eb, cc, steps, flow=flow,
                     kernel=kernel, u0=u0, pt0=pt0)
        got = _solve("asets_binary_dev_noalloc", d, shape, its, eb, cc, steps, flow=flow,
                     kernel=kernel, u0=u0, pt0=pt0)
        assert got[2] == ref[2]
        if eb > 0:
            assert ref[2] < its   # the stop test fired
        assert _same(got[0], ref[0]), f"case {k}: u differs"
        assert _same(got[1], ref[1]), f"case {k}: cvg differs"


def test_noalloc_rejections(cuda):
    shape = (256, 256, 256)   # 48 B per voxel of fp32 fused state: beyond the array
    N = 256 ** 3
    t = torch.ones(N, dtype=torch.float32, device=cuda)
    with pytest.raises(nat.AsetsError) as e:
        _solve("asets_binary_dev_noalloc", [t, t, t], shape, 5, 0.0, 0.35, 0.11,
               kernel="fused")
    assert e.value.status == nat.ASETS_ETOOBIG
    del t
    small = [torch.ones(8 * 6 * 4, dtype=torch.float32, device=cuda) for _ in range(3)]
    kms = C.c_double(0.0)
    opts = nat.make_opts(None, nat.FLOW_F32, nat.KERNEL_AUTO, 0, 0, 0, kms)
    p = nat.make_params(3, 0.0, 0.35, 0.11)
    out = torch.empty(8 * 6 * 4, dtype=torch.float32, device=cuda)
    cvg = torch.empty(3, dtype=torch.float32, device=cuda)
    itn = torch.empty(1, dtype=torch.int32, device=cuda)
    rc = nat.lib().asets_binary_dev_noalloc(3, *(x.data_ptr() for x in small), 8, 6, 4,
                                            C.byref(p), None, None, out.data_ptr(),
                                            cvg.data_ptr(), itn.data_ptr(), None,
                                            C.byref(opts))
    assert rc == nat.ASETS_EBADARG
    # alpha <= 0 is reported as by asets_binary_dev
    Cs, Ct, a = workloads.random_smooth_binary((8, 6, 4), seed=1, alpha=0.1)
    a[3, 2, 1] = 0.0
    d = [solver.to_device(x, cuda) for x in (Cs, Ct, a)]
    with pytest.raises(nat.AsetsError) as e:
        _solve("asets_binary_dev_noalloc", d, (8, 6, 4), 5, 0.0, 0.35, 0.11)
    assert e.value.status == nat.ASETS_EALPHA
