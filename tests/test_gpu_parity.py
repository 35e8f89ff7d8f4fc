"""GPU parity: the HIP kernels against the CPU oracle (reference operation order).

Bar: u bit-identical to the oracle's FP32 reference-order run at a fixed iteration count
(stronger than the max|u - u_ref| <= 1e-4 tolerance); cvg within 1e-6 relative of the
oracle's FP64-summed trace (the GPU sums |e| in FP64 in a different order); it_num equal
under the same stop rule.
"""
import numpy as np
import pytest
import torch

import oracle as orc
from asetsmatlabmaxflow_amd import mex, solver, workloads

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.asarray(a, dtype=np.float32).ravel(order="F")
    b = np.asarray(b, dtype=np.float32).ravel(order="F")
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _report(a, b):
    a = np.asarray(a, dtype=np.float64).ravel(order="F")
    b = np.asarray(b, dtype=np.float64).ravel(order="F")
    d = np.abs(a - b)
    return f"max|diff|={d.max():.3e} at {int(d.argmax())}, n_diff={int((a != b).sum())}"


def _run_binary(Cs, Ct, a, shape, maxit, eb, cc, steps, **kw):
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    return solver.binary(*d, shape, maxit, eb, cc, steps, **kw)


SHAPES_3D = [(1, 1, 1), (2, 3, 4), (5, 7, 3), (17, 13, 9), (64, 16, 20), (70, 33, 41),
             (130, 40, 37)]
SHAPES_2D = [(1, 1), (17, 13), (64, 64), (130, 77), (256, 256)]


@pytest.mark.parametrize("kernel", ["fused", "passes"])
@pytest.mark.parametrize("shape", SHAPES_3D)
def test_binary3d_bitexact(cuda, kernel, shape):
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=11 + sum(shape), alpha=0.1)
    maxit = 40
    ref = orc.binary(Cs, Ct, a, shape, maxit, 0.0, 0.25, 0.11)
    r = _run_binary(Cs, Ct, a, shape, maxit, 0.0, 0.25, 0.11, kernel=kernel)
    u = r.u.cpu().numpy()
    # a volume that starts at its fixed point (e.g. 1 voxel) has cvg[0] == 0 <= 0 and
    # stops at iteration 0, as the reference loop does
    assert r.it_num == ref.it_num
    assert _bits_equal(u, ref.u), _report(u, ref.u)
    np.testing.assert_allclose(r.cvg.cpu().numpy(), ref.cvg64.astype(np.float32), rtol=1e-6,
                               atol=0)


@pytest.mark.parametrize("kernel", ["fused", "passes"])
@pytest.mark.parametrize("shape", SHAPES_2D)
def test_binary2d_bitexact(cuda, kernel, shape):
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=7 + sum(shape), alpha=0.25)
    maxit = 60
    ref = orc.binary(Cs, Ct, a, shape, maxit, 0.0, 0.2, 0.16)
    r = _run_binary(Cs, Ct, a, shape, maxit, 0.0, 0.2, 0.16, kernel=kernel)
    u = r.u.cpu().numpy()
    assert _bits_equal(u, ref.u), _report(u, ref.u)
    np.testing.assert_allclose(r.cvg.cpu().numpy(), ref.cvg64.astype(np.float32), rtol=1e-6,
                               atol=0)


@pytest.mark.parametrize("tile_x,zchunk", [(8, 1), (8, 5), (16, 3), (16, 64), (32, 2),
                                           (32, 17), (8, 23), (16, 7)])
def test_fused_tilings_bitexact(cuda, tile_x, zchunk):
    shape = (70, 45, 23)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=5, count=4)
    maxit = 30
    ref = orc.binary(Cs, Ct, a, shape, maxit, 0.0, workloads.BINARY3D_CC,
                     workloads.BINARY3D_STEPS)
    r = _run_binary(Cs, Ct, a, shape, maxit, 0.0, workloads.BINARY3D_CC,
                    workloads.BINARY3D_STEPS, kernel="fused", tile_x=tile_x, zchunk=zchunk)
    u = r.u.cpu().numpy()
    assert _bits_equal(u, ref.u), _report(u, ref.u)


def test_binary3d_contrast_alpha_long_run(cuda):
    """The 3D workload generator at reduced size, 150 iterations."""
    shape = (96, 80, 72)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=1002, count=6)
    maxit = 150
    ref = orc.binary(Cs, Ct, a, shape, maxit, 0.0, workloads.BINARY3D_CC,
                     workloads.BINARY3D_STEPS)
    for kernel in ("fused", "passes"):
        r = _run_binary(Cs, Ct, a, shape, maxit, 0.0, workloads.BINARY3D_CC,
                        workloads.BINARY3D_STEPS, kernel=kernel)
        u = r.u.cpu().numpy()
        assert _bits_equal(u, ref.u), (kernel, _report(u, ref.u))


@pytest.mark.parametrize("kernel", ["fused", "passes"])
def test_stop_rule_matches(cuda, kernel):
    """errBound > 0: stop on the first iteration with cvg <= errBound (<=, 0-based it_num,
    asetsBinaryMF3D_mex.c:183-192); entries after the stop index stay 0."""
    shape = (33, 29, 17)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=3, alpha=0.2)
    full = orc.binary(Cs, Ct, a, shape, 200, 0.0, 0.25, 0.11, stop_rule=1)
    k = 57
    eb = float(np.float32(full.cvg64[k]))
    ref = orc.binary(Cs, Ct, a, shape, 200, eb, 0.25, 0.11, stop_rule=1)
    r = _run_binary(Cs, Ct, a, shape, 200, eb, 0.25, 0.11, kernel=kernel, poll_every=7)
    assert r.it_num == ref.it_num <= k
    cvg = r.cvg.cpu().numpy()
    assert np.all(cvg[r.it_num + 1:] == 0)
    assert _bits_equal(r.u.cpu().numpy(), ref.u)


def _first_new_minimum(cvg64, lo):
    """An iteration k >= lo whose cvg is below every earlier one, and an errBound halfway
    between, so that the first iteration with cvg <= errBound is k by a wide margin."""
    for k in range(lo, len(cvg64)):
        prev = cvg64[:k].min()
        if cvg64[k] < prev and (prev - cvg64[k]) > 1e-4 * prev:
            return k, float(np.float32(0.5 * (cvg64[k] + prev)))
    raise AssertionError("no new minimum of cvg")


@pytest.mark.parametrize("flow", ["f32", "f16"])
@pytest.mark.parametrize("shape", [(64, 29, 17), (36, 20, 9), (33, 29, 17), (128, 40)])
def test_stop_found_by_next_launch(cuda, flow, shape):
    """errBound > 0 on the fused kernel: every workgroup of the launch after iteration k
    runs k's stop test, and when it fires that launch writes u out of the final state
    (ny a multiple of 4, u asked for alone; else the copy after the loop does).  The stop
    falls inside a poll batch, on a batch's last launch, on the solve's last iteration
    (the test after the loop finds it) and one before (the last launch finds it); with
    and without the state outputs.  fp32 against the oracle; fp16 against its own run
    of the same iteration count with errBound 0."""
    cc, steps = 0.25, 0.11
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=21 + sum(shape), alpha=0.2)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    if flow == "f32":
        trace = orc.binary(Cs, Ct, a, shape, 120, 0.0, cc, steps, stop_rule=1).cvg64
    else:
        trace = solver.binary(*d, shape, 120, 0.0, cc, steps, kernel="fused",
                              flow=flow).cvg.cpu().numpy().astype(np.float64)
    k, eb = _first_new_minimum(trace, 20)
    cases = [(120, 7, False), (120, k + 1, False), (k + 1, 0, False), (k + 2, 0, False),
             (120, 7, True), (k + 2, 0, True)]
    for maxit, poll, state in cases:
        print(f"stop case shape={shape} flow={flow} k={k} maxit={maxit} poll={poll} "
              f"state={state}", flush=True)
        r = solver.binary(*d, shape, maxit, eb, cc, steps, kernel="fused", flow=flow,
                          poll_every=poll, want_state=state)
        assert r.it_num == k, (r.it_num, k)
        cvg = r.cvg.cpu().numpy()
        assert np.all(cvg[k + 1:] == 0)
        u = r.u.cpu().numpy()
        if flow == "f32":
            ref = orc.binary(Cs, Ct, a, shape, maxit, eb, cc, steps, stop_rule=1)
            assert ref.it_num == k
            assert _bits_equal(u, ref.u), _report(u, ref.u)
            np.testing.assert_allclose(cvg[:k + 1], ref.cvg64[:k + 1].astype(np.float32),
                                       rtol=1e-6, atol=0)
        else:
            fixed = solver.binary(*d, shape, k + 1, 0.0, cc, steps, kernel="fused", flow=flow)
            assert _bits_equal(u, fixed.u.cpu().numpy())
            assert np.array_equal(cvg[:k + 1], fixed.cvg.cpu().numpy())


@pytest.mark.parametrize("kernel", ["fused", "passes"])
def test_warm_start(cuda, kernel):
    shape = (40, 21, 13)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=9, alpha=0.15)
    rng = np.random.default_rng(0)
    u0 = rng.random(shape).astype(np.float32)
    pt0 = np.minimum(Cs, Ct) + rng.normal(0, 0.01, shape).astype(np.float32)
    ref = orc.binary(Cs, Ct, a, shape, 25, 0.0, 0.25, 0.11, u0=u0, pt0=pt0)
    d = [solver.to_device(x) for x in (Cs, Ct, a, u0, pt0)]
    r = solver.binary(d[0], d[1], d[2], shape, 25, 0.0, 0.25, 0.11, u0=d[3], pt0=d[4],
                      kernel=kernel)
    assert _bits_equal(r.u.cpu().numpy(), ref.u)


@pytest.mark.parametrize("kernel", ["fused", "passes"])
def test_final_state_matches(cuda, kernel):
    shape = (19, 11, 7)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=4, alpha=0.1)
    ref = orc.binary(Cs, Ct, a, shape, 12, 0.0, 0.25, 0.11, want_state=True)
    r = _run_binary(Cs, Ct, a, shape, 12, 0.0, 0.25, 0.11, kernel=kernel, want_state=True)
    for k in ("ps", "pt", "bx", "by", "bz"):
        assert _bits_equal(r.state[k].cpu().numpy(), ref.state[k]), k


@pytest.mark.parametrize("kernel", ["fused", "passes"])
@pytest.mark.parametrize("shape,nlab", [((1, 1, 1), 2), ((17, 13, 9), 3), ((33, 20, 11), 6),
                                        ((70, 9, 5), 10), ((65, 18, 12), 7)])
def test_potts3d_bitexact(cuda, shape, nlab, kernel):
    Ct, a = workloads.random_smooth_potts(shape, nlab, seed=sum(shape) + nlab)
    ref = orc.potts(Ct, a, shape, nlab, 30, 0.0, 0.25, 0.11)
    d = [solver.to_device(x) for x in (Ct, a)]
    r = solver.potts(*d, shape, nlab, 30, 0.0, 0.25, 0.11, kernel=kernel)
    u = r.u.cpu().numpy()
    assert _bits_equal(u, ref.u), _report(u, ref.u)
    np.testing.assert_allclose(r.cvg.cpu().numpy(), ref.cvg64.astype(np.float32), rtol=1e-6)


@pytest.mark.parametrize("kernel", ["fused", "passes"])
@pytest.mark.parametrize("shape,nlab", [((17, 13), 3), ((64, 48), 4), ((100, 77), 10)])
def test_potts2d_bitexact(cuda, shape, nlab, kernel):
    Ct, a = workloads.random_smooth_potts(shape, nlab, seed=sum(shape), alpha_step=0.05)
    ref = orc.potts(Ct, a, shape, nlab, 40, 0.0, 0.2, 0.16)
    d = [solver.to_device(x) for x in (Ct, a)]
    r = solver.potts(*d, shape, nlab, 40, 0.0, 0.2, 0.16, kernel=kernel)
    assert _bits_equal(r.u.cpu().numpy(), ref.u), _report(r.u.cpu().numpy(), ref.u)


# ---- the mex-ABI boundary -------------------------------------------------------------

def test_mex_binary3d(cuda):
    mex.set_quiet(True)
    shape = (21, 17, 9)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=21, alpha=0.1)
    pars = np.array([21, 17, 9, 35, 0.0, 0.25, 0.11], dtype=np.float32)
    u, cvg, it, t = mex.asetsBinaryMF3D_mex(Cs, Ct, a, pars)
    ref = orc.binary(Cs, Ct, a, shape, 35, 0.0, 0.25, 0.11)
    assert u.dtype == np.float32 and u.shape == shape
    assert cvg.dtype == np.float32 and cvg.shape == (35, 1)
    assert it.dtype == np.uint16 and it.shape == (1, 1) and int(it[0, 0]) == 35
    assert t.dtype == np.float32 and t.shape == (1, 1) and t[0, 0] > 0
    assert _bits_equal(u, ref.u)


def test_mex_fewer_outputs(cuda):
    """[u] = asetsBinaryMF3D_mex(...) and a call with no output: erriter is a temporary,
    u is the same, and no mxArray is left behind."""
    mex.set_quiet(True)
    shape = (21, 17, 9)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=21, alpha=0.1)
    pars = np.array([21, 17, 9, 35, 0.0, 0.25, 0.11], dtype=np.float32)
    base = mex.live_arrays()
    (u,) = mex.call("asetsBinaryMF3D_mex", Cs, Ct, a, pars, nlhs=1)
    assert mex.call("asetsBinaryMF3D_mex", Cs, Ct, a, pars, nlhs=0) == []
    ref = orc.binary(Cs, Ct, a, shape, 35, 0.0, 0.25, 0.11)
    assert _bits_equal(u, ref.u)
    assert mex.live_arrays() == base


def test_mex_binary2d_config1(cuda):
    """256x256 ellipse image with the pars of the 2D workload, 300 fixed iterations."""
    mex.set_quiet(True)
    Cs, Ct, a, pars, _ = workloads.binary2d_ellipses()
    u, cvg, it, t = mex.asetsBinaryMF2D_mex(Cs, Ct, a, pars)
    ref = orc.binary(Cs, Ct, a, (256, 256), 300, 0.0, 0.3, 0.16)
    assert u.shape == (256, 256) and int(it[0, 0]) == ref.it_num
    assert _bits_equal(u, ref.u), _report(u, ref.u)
    lab, lab_ref = u > 0.5, ref.u.reshape(256, 256, order="F") > 0.5
    assert (lab == lab_ref).mean() == 1.0


def test_mex_potts(cuda):
    mex.set_quiet(True)
    shape, nlab = (15, 12, 6), 4
    Ct, a = workloads.random_smooth_potts(shape, nlab, seed=2)
    pars = np.array([15, 12, 6, nlab, 20, 0.0, 0.25, 0.11], dtype=np.float32)
    u, cvg, it, t = mex.asetsPotts3D_mex(Ct, a, pars)
    ref = orc.potts(Ct, a, shape, nlab, 20, 0.0, 0.25, 0.11)
    assert u.shape == shape + (nlab,)
    assert _bits_equal(u, ref.u)
    Ct2, a2 = workloads.random_smooth_potts((15, 12), 3, seed=3)
    pars2 = np.array([15, 12, 3, 20, 0.0, 0.2, 0.16], dtype=np.float32)
    u2, _, _, _ = mex.asetsPotts2D_mex(Ct2, a2, pars2)
    ref2 = orc.potts(Ct2, a2, (15, 12), 3, 20, 0.0, 0.2, 0.16)
    assert u2.shape == (15, 12, 3) and _bits_equal(u2, ref2.u)


def test_mex_errors_and_edges(cuda):
    mex.set_quiet(True)
    shape = (8, 6, 4)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=1, alpha=0.1)
    bad = a.copy()
    bad[3, 2, 1] = 0.0
    pars = np.array([8, 6, 4, 10, 0.0, 0.25, 0.11], dtype=np.float32)
    with pytest.raises(mex.MexError, match="alpha"):
        mex.asetsBinaryMF3D_mex(Cs, Ct, bad, pars)
    # maxIt = 0: no iteration, u is the initial cut, numIter 0, no alpha check
    pars0 = pars.copy()
    pars0[3] = 0
    u, cvg, it, _ = mex.asetsBinaryMF3D_mex(Cs, Ct, bad, pars0)
    assert int(it[0, 0]) == 0 and cvg.shape == (0, 1)
    np.testing.assert_array_equal(u, (Cs - Ct >= 0).astype(np.float32))
    # the library is still usable after an error
    u, _, it, _ = mex.asetsBinaryMF3D_mex(Cs, Ct, a, pars)
    assert int(it[0, 0]) == 10


def test_fused_planar_fallback(cuda):
    """A volume whose block-interleaved state would pass 2^31 bytes per buffer runs the
    same fused kernel on planar fields: a 4352 x 4120 plane ((1 + 4) planes x 4120 columns
    x 68 blocks x 1536 B > 2^31), 5 iterations, bit for bit with the oracle."""
    shape = (4352, 4120)
    nyb = -(-shape[0] // 64)
    assert 5 * shape[1] * nyb * 1536 >= 2 ** 31
    print(f"planar fallback {shape}", flush=True)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=41, alpha=0.25)
    ref = orc.binary(Cs, Ct, a, shape, 5, 0.0, 0.2, 0.16)
    r = _run_binary(Cs, Ct, a, shape, 5, 0.0, 0.2, 0.16, kernel="fused")
    u = r.u.cpu().numpy()
    assert _bits_equal(u, ref.u), _report(u, ref.u)


def test_fused_equals_passes_full_size(cuda):
    """256^3 workload: the fused kernel and the per-pass kernels agree bit for bit over
    50 iterations (two independent GPU implementations of the same operation order)."""
    shape = (256, 256, 256)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=1002)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    del Cs, Ct, a
    rf = solver.binary(*d, shape, 50, 0.0, workloads.BINARY3D_CC, workloads.BINARY3D_STEPS,
                       kernel="fused")
    rp = solver.binary(*d, shape, 50, 0.0, workloads.BINARY3D_CC, workloads.BINARY3D_STEPS,
                       kernel="passes")
    assert torch.equal(rf.u.view(torch.int32), rp.u.view(torch.int32))
    torch.testing.assert_close(rf.cvg, rp.cvg, rtol=1e-6, atol=0)


@pytest.mark.parametrize("mode,d", [(0, 0.35), (0, 0.25), (0, 0.3), (0, 0.2), (0, 1.0 / 3.0),
                                    (0, 0.11), (0, 7.0), (0, 1.7e-3), (1, 1.0)])
def test_short_division_is_ieee(cuda, mode, d):
    """The kernels' short division form equals IEEE division for every float numerator
    (mode 0: x/d) and every reciprocal in its range (mode 1: 1/x): 2^32 cases each."""
    import ctypes as C

    from asetsmatlabmaxflow_amd import _native as nat

    bad = C.c_uint64(0)
    nat.check(nat.lib().asets_selftest_division(mode, float(np.float32(d)), C.byref(bad)))
    assert bad.value == 0


# ---- fp16 flow storage (fp32 arithmetic) ----------------------------------------------

@pytest.mark.parametrize("shape,iters", [((96, 80, 72), 150), ((130, 77), 300)])
def test_fp16_flow_storage_error(cuda, shape, iters):
    """ps, pt, bx, by, bz stored in fp16, u and all arithmetic in fp32.  Reported against
    the FP64 oracle next to the fp32 kernel's own error; bar: labels (u > 0.5) agree on
    >= 99.9 % of voxels and max |u - u_f64| <= 0.05."""
    if len(shape) == 3:
        Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=1002, count=6)
        cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    else:
        Cs, Ct, a, _, _ = workloads.binary2d_ellipses(n=shape[0], seed=1001)[:4] + (None,)
        Cs, Ct, a = Cs[:, :shape[1]], Ct[:, :shape[1]], a[:, :shape[1]]
        cc, st = 0.3, 0.16
    ref64 = orc.binary(Cs, Ct, a, shape, iters, 0.0, cc, st, precision="f64")
    r16 = _run_binary(Cs, Ct, a, shape, iters, 0.0, cc, st, kernel="fused", flow="f16")
    r32 = _run_binary(Cs, Ct, a, shape, iters, 0.0, cc, st, kernel="fused")
    u16, u32 = r16.u.cpu().numpy().astype(np.float64), r32.u.cpu().numpy().astype(np.float64)
    err16 = np.abs(u16 - ref64.u).max()
    err32 = np.abs(u32 - ref64.u).max()
    agree = ((u16 > 0.5) == (ref64.u > 0.5)).mean()
    print(f"fp16 storage {shape} x{iters}: max|u-u64| fp16 {err16:.3e} fp32 {err32:.3e}, "
          f"label agreement {agree:.6f}")
    assert err32 < 1e-4
    assert err16 < 0.05 and agree >= 0.999


def test_fp16_requires_fused(cuda):
    from asetsmatlabmaxflow_amd import _native as nat

    shape = (8, 8, 4)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=1)
    with pytest.raises(nat.AsetsError, match="fused"):
        _run_binary(Cs, Ct, a, shape, 3, 0.0, 0.25, 0.11, kernel="passes", flow="f16")


# ---- batched uneven volumes --------------------------------------------------------------

BATCH_SHAPES = [(40, 33, 21), (64, 64, 17), (17, 90, 33), (70, 45, 23), (33, 20, 66),
                (130, 8, 9)]


def _batch_inputs():
    vols = []
    for i, sh in enumerate(BATCH_SHAPES):
        Cs, Ct, a, _ = workloads.binary3d_ellipsoids(sh, seed=1004000 + i, count=3)
        vols.append((Cs, Ct, a, sh))
    return vols


@pytest.mark.parametrize("flow", ["f32", "f16"])
def test_batched_matches_single(cuda, flow, tmp_path, monkeypatch):
    """Each volume of a batched solve (autotuned tiling, graph replay, per-volume stop)
    equals its own single-volume solve bit for bit, including the stop iteration."""
    monkeypatch.setenv("ASETS_TUNE_CACHE", str(tmp_path / "tune.txt"))
    host = _batch_inputs()
    dev = [tuple(solver.to_device(x) for x in v[:3]) + (v[3],) for v in host]
    eb = 2e-4
    res, secs, _ = solver.binary_batched(dev, 300, eb, 0.35, 0.11, flow=flow)
    its = []
    for (Cs, Ct, a, sh), r in zip(dev, res):
        single = solver.binary(Cs, Ct, a, sh, 300, eb, 0.35, 0.11, kernel="fused", flow=flow)
        assert r.it_num == single.it_num, sh
        assert torch.equal(r.u.view(torch.int32), single.u.view(torch.int32)), sh
        torch.testing.assert_close(r.cvg[:r.iterations], single.cvg[:single.iterations],
                                   rtol=1e-6, atol=0)
        its.append(r.it_num)
    assert len(set(its)) > 1   # volumes stopped at different iterations
    lines = (tmp_path / "tune.txt").read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("batched3d:")
    # second call with the same shape bucket: cache hit, no new tuning entry
    solver.binary_batched(dev, 5, 0.0, 0.35, 0.11, flow=flow)
    assert len((tmp_path / "tune.txt").read_text().splitlines()) == 1


@pytest.mark.parametrize("flow", ["f32", "f16"])
def test_batched_workload_shapes_to_1e4(cuda, flow, tmp_path, monkeypatch):
    """The batched workload's first 8 volumes (its generator and seeds: axes uniform in
    {64..192}, volume v from seed 1004*1000+v) in one batched call to errBound 1e-4: each
    equals its own single-volume solve bit for bit, stop iteration included; against the
    FP64 oracle run for the same iteration count on the two smallest volumes, max |u -
    u_f64| stays below 1e-5 with fp32 flows (measured 7.8e-7) and 5e-3 with fp16 flows
    (measured 1.4e-3), with at least 99.99 % of the labels u > 0.5 equal."""
    monkeypatch.setenv("ASETS_TUNE_CACHE", str(tmp_path / "tune.txt"))
    rng = np.random.default_rng(1004)
    sizes = [tuple(int(v) for v in rng.integers(64, 193, 3)) for _ in range(8)]
    print(f"batched workload shapes {sizes} flow={flow}", flush=True)
    host = []
    for v, sh in enumerate(sizes):
        Cs, Ct, a, _ = workloads.binary3d_ellipsoids(sh, seed=1004 * 1000 + v)
        host.append((Cs, Ct, a, sh))
    dev = [tuple(solver.to_device(x) for x in h[:3]) + (h[3],) for h in host]
    cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    res, _, _ = solver.binary_batched(dev, 2000, 1e-4, cc, st, flow=flow)
    for (Cs, Ct, a, sh), r in zip(dev, res):
        single = solver.binary(Cs, Ct, a, sh, 2000, 1e-4, cc, st, kernel="fused", flow=flow)
        assert r.it_num == single.it_num < 2000, sh
        assert torch.equal(r.u.view(torch.int32), single.u.view(torch.int32)), sh
    bound = 1e-5 if flow == "f32" else 5e-3
    for v in sorted(range(len(sizes)), key=lambda i: int(np.prod(sizes[i])))[:2]:
        Cs, Ct, a, sh = host[v]
        r64 = orc.binary(Cs, Ct, a, sh, res[v].iterations, 0.0, cc, st, precision="f64")
        u = res[v].u.cpu().numpy().astype(np.float64)
        assert np.abs(u - r64.u).max() < bound, sh
        assert np.mean((u > 0.5) == (r64.u > 0.5)) >= 0.9999, sh


def test_batched_vs_oracle(cuda):
    """Fixed iteration count: batched u equals the CPU oracle bit for bit."""
    host = _batch_inputs()[:3]
    dev = [tuple(solver.to_device(x) for x in v[:3]) + (v[3],) for v in host]
    res, _, _ = solver.binary_batched(dev, 25, 0.0, 0.35, 0.11, tile_x=8, zchunk=8)
    for (Cs, Ct, a, sh), r in zip(host, res):
        ref = orc.binary(Cs, Ct, a, sh, 25, 0.0, 0.35, 0.11)
        assert _bits_equal(r.u.cpu().numpy(), ref.u), sh


@pytest.mark.parametrize("flow", ["f32", "f16"])
@pytest.mark.parametrize("shape", [(37, 29, 21), (70, 45)])
def test_cold_start_iteration_equals_initial_state(cuda, flow, shape):
    """A cold start forms the initial cut (u = Cs - Ct >= 0, ps = pt = min(Cs, Ct) as the
    flow storage rounds it, flows 0; asetsBinaryMF3D_mex.c:215-219) inside its first
    fused iteration.  The same solve warm-started from that state (written by a separate
    kernel, first iteration reading it back) gives the same bits: u, cvg and the stop
    iteration, at fixed counts and under a stop rule."""
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=11, alpha=0.2)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    u0 = solver.to_device((Cs - Ct >= 0).astype(np.float32))
    pt0 = solver.to_device(np.where(Cs < Ct, Cs, Ct).astype(np.float32))
    for mi, eb in ((1, 0.0), (7, 0.0), (300, 2e-3)):
        rc = solver.binary(*d, shape, mi, eb, 0.3, 0.16, kernel="fused", flow=flow)
        rw = solver.binary(*d, shape, mi, eb, 0.3, 0.16, kernel="fused", flow=flow,
                           u0=u0, pt0=pt0)
        assert rc.it_num == rw.it_num, (mi, eb)
        assert torch.equal(rc.u.view(torch.int32), rw.u.view(torch.int32)), (mi, eb)
        assert torch.equal(rc.cvg.view(torch.int32), rw.cvg.view(torch.int32)), (mi, eb)
    if flow == "f32":   # and the oracle, one iteration
        ref = orc.binary(Cs, Ct, a, shape, 1, 0.0, 0.3, 0.16)
        r1 = solver.binary(*d, shape, 1, 0.0, 0.3, 0.16, kernel="fused")
        assert _bits_equal(r1.u.cpu().numpy(), ref.u)


def test_batched_zero_and_one_iteration(cuda):
    """Batched solve with maxIter 0 (u is the initial cut, it_num 0) and 1 (the
    cold-start iteration alone) against the oracle."""
    host = _batch_inputs()[:3]
    dev = [tuple(solver.to_device(x) for x in v[:3]) + (v[3],) for v in host]
    r0, _, _ = solver.binary_batched(dev, 0, 0.0, 0.35, 0.11)
    r1, _, _ = solver.binary_batched(dev, 1, 0.0, 0.35, 0.11)
    for (Cs, Ct, a, sh), a0, a1 in zip(host, r0, r1):
        assert a0.it_num == 0
        np.testing.assert_array_equal(a0.u.cpu().numpy(),
                                      (Cs - Ct >= 0).astype(np.float32).ravel(order="F"))
        ref = orc.binary(Cs, Ct, a, sh, 1, 0.0, 0.35, 0.11)
        assert a1.it_num == 1 and _bits_equal(a1.u.cpu().numpy(), ref.u), sh


@pytest.mark.parametrize("flow", ["f32", "f16"])
@pytest.mark.parametrize("shape,tile_x,zchunk", [((70, 45, 33), 8, 0), ((70, 45, 33), 16, 0),
                                                 ((70, 45, 33), 32, 0), ((130, 77), 16, 0),
                                                 ((200, 96, 96), 8, 1)])
def test_deferred_stop_test_same_bits(cuda, flow, shape, tile_x, zchunk):
    """errBound 0 folds each iteration's stop test into the end of the next iteration's
    first workgroup (workgroups of 256, 512 or 1024 threads playing the finalize kernel's
    1024).  A negative errBound (never met) runs it at the start of every workgroup of
    the next iteration's kernel, or, past 2048 workgroups (4608 here with 64 x 8 tiles
    of one plane), in the separate finalize kernel.  The cvg traces and u are the same
    bits."""
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=23, alpha=0.15)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    kw = dict(kernel="fused", flow=flow, tile_x=tile_x, zchunk=zchunk)
    r0 = solver.binary(*d, shape, 25, 0.0, 0.3, 0.14, **kw)
    rn = solver.binary(*d, shape, 25, -1.0, 0.3, 0.14, **kw)
    assert r0.it_num == rn.it_num == 25
    assert torch.equal(r0.cvg.view(torch.int32), rn.cvg.view(torch.int32))
    assert torch.equal(r0.u.view(torch.int32), rn.u.view(torch.int32))
    assert bool((r0.cvg > 0).all())


@pytest.mark.parametrize("shape,nlab", [((33, 20, 11), 6), ((70, 9, 5), 4), ((65, 18, 12), 3),
                                        ((64, 48), 4)])
def test_potts_deferred_stop_test_same_bits(cuda, shape, nlab):
    """Fused Potts: errBound 0 runs each iteration's stop test in the first workgroup of
    the next flow launch; a negative errBound (never met) keeps the finalize kernel
    after every iteration.  u, cvg and the iteration count are the same bits."""
    Ct, a = workloads.random_smooth_potts(shape, nlab, seed=sum(shape) + 3 * nlab)
    d = [solver.to_device(x) for x in (Ct, a)]
    r0 = solver.potts(*d, shape, nlab, 25, 0.0, 0.25, 0.11, kernel="fused")
    rn = solver.potts(*d, shape, nlab, 25, -1.0, 0.25, 0.11, kernel="fused")
    assert r0.it_num == rn.it_num == 25
    assert torch.equal(r0.cvg.view(torch.int32), rn.cvg.view(torch.int32))
    assert torch.equal(r0.u.view(torch.int32), rn.u.view(torch.int32))
    assert bool((r0.cvg > 0).all())


def test_alpha_error_every_solve(cuda):
    """alpha <= 0 anywhere (asetsBinaryMF3D_mex.c:317-319): the binary solves (fused fp32
    and fp16, per-pass; cold, warm and resumed), Potts and the batched solve report it,
    the batched one naming
    the volume.  The check runs on the device before the first iteration and is read
    with the first stop-flag poll; an error leaves the library usable, and a following
    solve equals the oracle."""
    from asetsmatlabmaxflow_amd import _native as nat

    shape = (8, 6, 4)
    Cs, Ct, a = workloads.random_smooth_binary(shape, seed=1, alpha=0.1)
    bad = a.copy()
    bad[3, 2, 1] = -1.0
    g = [solver.to_device(x) for x in (Cs, Ct, a)]
    b = [solver.to_device(x) for x in (Cs, Ct, bad)]
    for kernel, flow, eb in (("fused", "f32", 1e-4), ("fused", "f16", 0.0),
                             ("passes", "f32", 0.0)):
        with pytest.raises(nat.AsetsError, match="alpha"):
            solver.binary(*b, shape, 40, eb, 0.25, 0.11, kernel=kernel, flow=flow)
    # warm starts and resumes (checked while the state is laid out, or by the per-pass
    # solver's own check)
    u0 = torch.full((int(np.prod(shape)),), 0.5, dtype=torch.float32, device=cuda)
    pt0 = torch.minimum(b[0], b[1])
    part = solver.binary(*g, shape, 3, 0.0, 0.25, 0.11, want_state=True)
    full_state = dict(part.state, u=part.u)
    paired = {k: full_state[k] for k in ("u", "bx", "by", "bz")}
    for kernel, flow in (("fused", "f32"), ("fused", "f16"), ("passes", "f32")):
        with pytest.raises(nat.AsetsError, match="alpha"):
            solver.binary(*b, shape, 40, 0.0, 0.25, 0.11, u0=u0, pt0=pt0, kernel=kernel,
                          flow=flow)
        for st in (full_state, paired):
            with pytest.raises(nat.AsetsError, match="alpha"):
                solver.binary_resume(*b, shape, st, 10, 1e-4, 0.25, 0.11, kernel=kernel,
                                     flow=flow)
    with pytest.raises(nat.AsetsError, match="volume 1: alpha"):
        solver.binary_batched([tuple(g) + (shape,), tuple(b) + (shape,)], 10, 0.0, 0.25, 0.11)
    Ctp, ap = workloads.random_smooth_potts((8, 6, 4), 3, seed=2)
    ap = ap.copy()
    ap[1, 1, 1, 2] = 0.0
    with pytest.raises(nat.AsetsError, match="alpha"):
        solver.potts(*(solver.to_device(x) for x in (Ctp, ap)), (8, 6, 4), 3, 10, 0.0, 0.2, 0.11)
    r = solver.binary(*g, shape, 10, 0.0, 0.25, 0.11)
    ref = orc.binary(Cs, Ct, a, shape, 10, 0.0, 0.25, 0.11)
    assert _bits_equal(r.u.cpu().numpy(), ref.u)
    rb, _, _ = solver.binary_batched([tuple(g) + (shape,)], 10, 0.0, 0.25, 0.11)
    assert rb[0].it_num == 10 and _bits_equal(rb[0].u.cpu().numpy(), ref.u)


def test_potts_config3_workload_vs_oracle(cuda):
    """The 192^3 Potts workload's generator (L = 6 smoothed Voronoi labels, alpha 0.8,
    cc 0.25, steps 0.11, 300 fixed iterations) at 64x56x48: u of the fused and the
    per-pass kernels equals the oracle's bit for bit (asetsPotts3D_mex.c:144-217,
    474-521), and cvg agrees with the oracle's FP64 sum."""
    shape, L, its = (64, 56, 48), 6, 300
    print(f"potts workload vs oracle {shape} L={L} {its} iterations", flush=True)
    Ct, a, _ = workloads.potts3d_voronoi(shape, nlab=L, seed=1003)
    cc, st = workloads.POTTS3D_CC, workloads.POTTS3D_STEPS
    ref = orc.potts(Ct, a, shape, L, its, 0.0, cc, st)
    d = [solver.to_device(x) for x in (Ct, a)]
    for kern in ("fused", "passes"):
        r = solver.potts(*d, shape, L, its, 0.0, cc, st, kernel=kern)
        assert r.it_num == ref.it_num == its
        u = r.u.cpu().numpy()
        assert _bits_equal(u, ref.u), (kern, _report(u, ref.u))
        np.testing.assert_allclose(r.cvg.cpu().numpy(), ref.cvg64, rtol=1e-6, atol=0)


def test_potts_config3_full_size_properties(cuda):
    """The 192^3, L = 6 Potts workload at full size, 300 fixed iterations: the labels sum
    to 1 at every voxel to round-off (the coupling of asetsPotts3D_mex.c:495-514 keeps
    sum_l u_l = 1 in exact arithmetic; the oracle's fp32 trajectory stays within 1.2e-6
    of it at 48x40x32), and the fused and per-pass kernels agree bit for bit."""
    shape, L, its = (192, 192, 192), 6, 300
    print(f"potts workload full size {shape} L={L}", flush=True)
    Ct, a, _ = workloads.potts3d_voronoi(shape, nlab=L, seed=1003)
    cc, st = workloads.POTTS3D_CC, workloads.POTTS3D_STEPS
    d = [solver.to_device(x) for x in (Ct, a)]
    del Ct, a
    rf = solver.potts(*d, shape, L, its, 0.0, cc, st, kernel="fused")
    s = rf.u.view(L, -1).double().sum(0)
    assert float((s - 1).abs().max()) < 1e-5
    rp = solver.potts(*d, shape, L, its, 0.0, cc, st, kernel="passes")
    assert torch.equal(rf.u.view(torch.int32), rp.u.view(torch.int32))
    assert torch.equal(rf.cvg, rp.cvg)


def test_potts_voronoi_fused_vs_passes_and_stop(cuda):
    """Potts workload generator at reduced size: fused == per-pass kernels bit for bit over
    100 iterations, and the stop rule fires on the same iteration."""
    shape, L = (96, 80, 40), 6
    Ct, a, _ = workloads.potts3d_voronoi(shape, nlab=L, seed=1003)
    d = [solver.to_device(x) for x in (Ct, a)]
    rf = solver.potts(*d, shape, L, 100, 0.0, 0.25, 0.11, kernel="fused")
    rp = solver.potts(*d, shape, L, 100, 0.0, 0.25, 0.11, kernel="passes")
    assert torch.equal(rf.u.view(torch.int32), rp.u.view(torch.int32))
    eb = float(rf.cvg[40].item())
    sf = solver.potts(*d, shape, L, 100, eb, 0.25, 0.11, kernel="fused", poll_every=9)
    sp = solver.potts(*d, shape, L, 100, eb, 0.25, 0.11, kernel="passes")
    assert sf.it_num == sp.it_num <= 40
    assert torch.equal(sf.u.view(torch.int32), sp.u.view(torch.int32))
