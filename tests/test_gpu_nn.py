"""Warm-start network kernels (libasets_nn.so) against the float64 CPU oracle
(oracle/nn_oracle.py): each operation on its own with tight tolerances (inputs are
bf16-exact, so the only differences are fp32 vs fp64 accumulation and the final bf16
rounding), then the whole U-Net's loss and parameter gradients."""
import numpy as np
import pytest
import torch

import nn_oracle as NO
from asetsmatlabmaxflow_amd import warmstart as W

pytestmark = pytest.mark.gpu

BF16_REL = 2.0 ** -8   # one bf16 rounding (half an ulp is 2^-9) with margin


def _bf(a: np.ndarray, device):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device).to(torch.bfloat16)


def _f(t) -> np.ndarray:
    return t.float().cpu().numpy().astype(np.float64)


def _p(t):
    return None if t is None else t.data_ptr()


def _assert_bf16_close(got, ref, scale=None):
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max() if scale is None else scale
    err = np.abs(got - ref)
    bound = BF16_REL * np.abs(ref) + 1e-5 * max(scale, 1e-30)
    bad = err > bound
    assert not bad.any(), f"{bad.sum()} of {bad.size} out of tolerance; max err {err.max():.3e}"


# rows narrower than 64 voxels run the gathering kernel, wider ones the LDS-staged one
SHAPES = [(2, 3, 5, 7), (1, 8, 8, 8), (1, 4, 6, 40), (1, 3, 4, 70), (2, 2, 3, 64), (1, 4, 4, 128)]
CHANNELS = [(8, 16), (16, 32), (48, 16), (32, 64), (96, 32), (128, 64), (16, 48), (32, 96),
            (64, 128)]


@pytest.mark.parametrize("geo", SHAPES)
@pytest.mark.parametrize("cin,cout", CHANNELS)
@pytest.mark.parametrize("relu", [0, 1])
def test_conv3_forward(cuda, geo, cin, cout, relu):
    n, d, h, w = geo
    P = n * d * h * w
    rng = np.random.default_rng(cin * 1000 + cout + P)
    x = NO.bf16_round(rng.normal(0, 1, (P, cin)))
    wt = NO.bf16_round(rng.normal(0, 0.2, 27 * cout * cin))
    b = rng.normal(0, 0.5, cout).astype(np.float32)
    xd, wd = _bf(x, cuda), _bf(wt, cuda)
    bd = torch.from_numpy(b).to(cuda)
    out = torch.empty(P * cout, dtype=torch.bfloat16, device=cuda)
    W._ck(W.nnlib().asets_nn_conv3(_p(xd), n, d, h, w, cin, _p(wd), _p(bd), cout, relu,
                                   _p(out), None))
    torch.cuda.synchronize()
    ref = NO.conv3(NO.to_ncdhw(x, n, d, h, w, cin), NO.weight_to_torch(wt, cin, cout),
                   torch.as_tensor(b, dtype=torch.float64), bool(relu))
    _assert_bf16_close(_f(out), NO.from_ncdhw(ref).numpy())


@pytest.mark.parametrize("geo", SHAPES[:2] + [(2, 4, 6, 8), (1, 3, 2, 66), (2, 2, 2, 64)])
@pytest.mark.parametrize("cin,cout", [(8, 16), (16, 16), (48, 16), (96, 32), (128, 64),
                                      (64, 64), (32, 64)])
def test_conv3_backward(cuda, geo, cin, cout):
    """Input gradient (mirrored-tap convolution) and weight/bias gradients vs autograd."""
    n, d, h, w = geo
    P = n * d * h * w
    rng = np.random.default_rng(7 * cin + cout + P)
    x = NO.bf16_round(rng.normal(0, 1, (P, cin)))
    wt = NO.bf16_round(rng.normal(0, 0.2, 27 * cout * cin))
    gy = NO.bf16_round(rng.normal(0, 1, (P, cout)))
    L = W.nnlib()
    xd, wd, gyd = _bf(x, cuda), _bf(wt, cuda), _bf(gy, cuda)
    X = NO.to_ncdhw(x, n, d, h, w, cin).clone().requires_grad_(True)
    Wt = torch.as_tensor(wt.astype(np.float64)).requires_grad_(True)
    B = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    y = NO.conv3(X, NO.weight_to_torch(Wt, cin, cout), B, False)
    y.backward(NO.to_ncdhw(gy, n, d, h, w, cout))
    # weight and bias gradients
    dw = torch.zeros(27 * cout * cin, dtype=torch.float32, device=cuda)
    db = torch.zeros(cout, dtype=torch.float32, device=cuda)
    nbytes = L.asets_nn_wgrad_scratch_bytes(n, d, h, w, cin, cout)
    scratch = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
    W._ck(L.asets_nn_conv3_wgrad(_p(xd), n, d, h, w, cin, _p(gyd), cout, _p(dw), _p(db), 0,
                                 _p(scratch), nbytes, None))
    torch.cuda.synchronize()
    rw, rb_ = Wt.grad.numpy(), B.grad.numpy()
    np.testing.assert_allclose(dw.cpu().numpy(), rw, rtol=0, atol=1e-5 * np.abs(rw).max())
    np.testing.assert_allclose(db.cpu().numpy(), rb_, rtol=0, atol=1e-5 * np.abs(rb_).max())
    # accumulate mode adds
    W._ck(L.asets_nn_conv3_wgrad(_p(xd), n, d, h, w, cin, _p(gyd), cout, _p(dw), _p(db), 1,
                                 _p(scratch), nbytes, None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(dw.cpu().numpy(), 2 * rw, rtol=0, atol=2e-5 * np.abs(rw).max())
    if cin == 8:
        return   # the input layer needs no input gradient (and 8 is no supported cout)
    wf = torch.empty(27 * cout * cin, dtype=torch.bfloat16, device=cuda)
    gx = torch.empty(P * cin, dtype=torch.bfloat16, device=cuda)
    W._ck(L.asets_nn_flip_transpose(_p(wd), _p(wf), cout, cin, None))
    W._ck(L.asets_nn_conv3(_p(gyd), n, d, h, w, cout, _p(wf), None, cin, 0, _p(gx), None))
    torch.cuda.synchronize()
    _assert_bf16_close(_f(gx), NO.from_ncdhw(X.grad).numpy())


@pytest.mark.parametrize("cin,cout", [(8, 16), (48, 16), (128, 64)])
def test_conv3_backward_many_chunks(cuda, cin, cout):
    """17,760 voxels: the weight gradient runs over 18 voxel chunks (the last one and its
    last 32-voxel slab partly filled) whose partials are added in order."""
    test_conv3_backward(cuda, (1, 20, 24, 37), cin, cout)


def test_relu_mask(cuda):
    rng = np.random.default_rng(3)
    g = NO.bf16_round(rng.normal(0, 1, 1000))
    y = NO.bf16_round(np.maximum(rng.normal(0, 1, 1000), 0))
    gd, yd = _bf(g, cuda), _bf(y, cuda)
    W._ck(W.nnlib().asets_nn_relu_mask(_p(gd), _p(yd), 1000, None))
    assert np.array_equal(_f(gd), np.where(y > 0, g, 0.0))


@pytest.mark.parametrize("geo,c", [((2, 4, 6, 8), 16), ((1, 8, 2, 4), 64)])
def test_pool_and_upsample(cuda, geo, c):
    n, d, h, w = geo
    P = n * d * h * w
    rng = np.random.default_rng(P + c)
    L = W.nnlib()
    x = NO.bf16_round(rng.normal(0, 1, (P, c)))
    xd = _bf(x, cuda)
    # pool forward
    pd = torch.empty(P // 8 * c, dtype=torch.bfloat16, device=cuda)
    W._ck(L.asets_nn_pool2(_p(xd), n, d, h, w, c, _p(pd), None))
    X = NO.to_ncdhw(x, n, d, h, w, c).clone().requires_grad_(True)
    pool = NO.pool2(X)
    torch.cuda.synchronize()
    _assert_bf16_close(_f(pd), NO.from_ncdhw(pool).detach().numpy())
    # pool backward (and accumulate)
    gp = NO.bf16_round(rng.normal(0, 1, (P // 8, c)))
    gpd = _bf(gp, cuda)
    pool.backward(NO.to_ncdhw(gp, n, d // 2, h // 2, w // 2, c))
    gx = torch.zeros(P * c, dtype=torch.bfloat16, device=cuda)
    W._ck(L.asets_nn_pool2_bwd(_p(gpd), n, d, h, w, c, _p(gx), 0, None))
    torch.cuda.synchronize()
    ref = NO.from_ncdhw(X.grad).numpy()
    _assert_bf16_close(_f(gx), ref)
    W._ck(L.asets_nn_pool2_bwd(_p(gpd), n, d, h, w, c, _p(gx), 1, None))
    torch.cuda.synchronize()
    _assert_bf16_close(_f(gx), 2 * ref)
    # upsample into a channel slice of a wider buffer, and its gradient
    cd, off = c + 16, 8
    xc = NO.bf16_round(rng.normal(0, 1, (P // 8, c)))
    xcd = _bf(xc, cuda)
    cat = torch.zeros(P * cd, dtype=torch.bfloat16, device=cuda)
    W._ck(L.asets_nn_up2(_p(xcd), n, d, h, w, c, _p(cat), cd, off, None))
    torch.cuda.synchronize()
    XC = NO.to_ncdhw(xc, n, d // 2, h // 2, w // 2, c).clone().requires_grad_(True)
    up = NO.up2(XC)
    got = _f(cat).reshape(P, cd)
    assert np.array_equal(got[:, off:off + c], NO.from_ncdhw(up).detach().numpy().reshape(P, c))
    assert not got[:, :off].any() and not got[:, off + c:].any()
    gcat = NO.bf16_round(rng.normal(0, 1, (P, cd)))
    gcd = _bf(gcat, cuda)
    up.backward(NO.to_ncdhw(np.ascontiguousarray(gcat[:, off:off + c]), n, d, h, w, c))
    gxc = torch.empty(P // 8 * c, dtype=torch.bfloat16, device=cuda)
    W._ck(L.asets_nn_up2_bwd(_p(gcd), cd, off, n, d, h, w, c, _p(gxc), None))
    torch.cuda.synchronize()
    _assert_bf16_close(_f(gxc), NO.from_ncdhw(XC.grad).numpy())


@pytest.mark.parametrize("geo,ca,cs,cout", [((2, 4, 6, 8), 64, 64, 64), ((1, 6, 4, 10), 64, 32, 32),
                                           ((2, 2, 8, 6), 32, 16, 16), ((1, 8, 8, 8), 16, 8, 16),
                                           ((1, 4, 2, 128), 32, 16, 16), ((2, 2, 4, 64), 64, 32, 32),
                                           ((1, 2, 2, 66), 64, 64, 64)])
def test_conv3_upcat_equals_materialised(cuda, geo, ca, cs, cout):
    """The decoder convolution gathering [up2(coarse), skip] in place gives the same bits
    as upsample + channel copy + convolution."""
    n, d, h, w = geo
    P, Pc, cin = n * d * h * w, n * d * h * w // 8, ca + cs
    rng = np.random.default_rng(P + ca + cs)
    coarse = _bf(NO.bf16_round(rng.normal(0, 1, (Pc, ca))), cuda)
    skip = _bf(NO.bf16_round(rng.normal(0, 1, (P, cs))), cuda)
    wt = _bf(NO.bf16_round(rng.normal(0, 0.2, 27 * cout * cin)), cuda)
    b = torch.from_numpy(rng.normal(0, 0.5, cout).astype(np.float32)).to(cuda)
    L = W.nnlib()
    cat = torch.empty(P * cin, dtype=torch.bfloat16, device=cuda)
    ref = torch.empty(P * cout, dtype=torch.bfloat16, device=cuda)
    out = torch.full((P * cout,), 7.0, dtype=torch.bfloat16, device=cuda)
    W._ck(L.asets_nn_up2(_p(coarse), n, d, h, w, ca, _p(cat), cin, 0, None))
    W._ck(L.asets_nn_chan_copy(_p(cat), cin, ca, _p(skip), cs, 0, cs, P, 0, None))
    W._ck(L.asets_nn_conv3(_p(cat), n, d, h, w, cin, _p(wt), _p(b), cout, 1, _p(ref), None))
    W._ck(L.asets_nn_conv3_upcat(_p(coarse), ca, _p(skip), n, d, h, w, cin, _p(wt), _p(b),
                                 cout, 1, _p(out), None))
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int16), ref.view(torch.int16))


@pytest.mark.parametrize("shape,n", [((16, 24, 8), 2), ((128, 16, 8), 1)])
def test_inference_forward_equals_training_forward(cuda, shape, n):
    """The inference forward (decoder inputs gathered in the convolution) and the training
    forward (decoder inputs materialised) predict the same bits, with rows narrower and
    wider than the LDS-staged kernel's 64 voxels."""
    Cs, Ct, A, _, _ = _net_inputs(shape, n, 5)
    net = W.WarmStartNet(cuda, seed=4)
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A)]
    u1, p1 = (t.clone() for t in net.forward(d[0], d[1], d[2], shape, n, train=True))
    u2, p2 = net.forward(d[0], d[1], d[2], shape, n, train=False)
    torch.cuda.synchronize()
    assert torch.equal(u1, u2) and torch.equal(p1, p2)


def test_chan_copy_and_features(cuda):
    rng = np.random.default_rng(11)
    L = W.nnlib()
    P = 777
    src = NO.bf16_round(rng.normal(0, 1, (P, 32)))
    dst = NO.bf16_round(rng.normal(0, 1, (P, 48)))
    sd, dd = _bf(src, cuda), _bf(dst, cuda)
    W._ck(L.asets_nn_chan_copy(_p(dd), 48, 16, _p(sd), 32, 8, 16, P, 0, None))
    torch.cuda.synchronize()
    want = dst.copy()
    want[:, 16:32] = src[:, 8:24]
    assert np.array_equal(_f(dd).reshape(P, 48), want)
    W._ck(L.asets_nn_chan_copy(_p(dd), 48, 16, _p(sd), 32, 8, 16, P, 1, None))
    torch.cuda.synchronize()
    want[:, 16:32] = NO.bf16_round(want[:, 16:32] + src[:, 8:24])
    assert np.array_equal(_f(dd).reshape(P, 48), want)
    Cs, Ct = rng.uniform(0, 30, P).astype(np.float32), rng.uniform(0, 30, P).astype(np.float32)
    a = rng.uniform(0.2, 1.0, P).astype(np.float32)
    feat = torch.empty(P * 8, dtype=torch.bfloat16, device=cuda)
    dev = [torch.from_numpy(v).to(cuda) for v in (Cs, Ct, a)]
    W._ck(L.asets_nn_features(*(_p(t) for t in dev), P, _p(feat), None))
    torch.cuda.synchronize()
    assert np.array_equal(_f(feat).reshape(P, 8), NO.features(Cs, Ct, a))


def test_head_forward_backward(cuda):
    rng = np.random.default_rng(5)
    L = W.nnlib()
    P = 1000
    x = NO.bf16_round(np.maximum(rng.normal(0, 1, (P, 16)), 0))
    w = rng.normal(0, 0.3, 32).astype(np.float32)
    b = np.array([0.4, -0.1], np.float32)
    Cs, Ct = rng.uniform(0, 30, P).astype(np.float32), rng.uniform(0, 30, P).astype(np.float32)
    tu, tpt = rng.uniform(0, 1, P).astype(np.float32), rng.uniform(0, 30, P).astype(np.float32)
    xd = _bf(x, cuda)
    wd, bd, Csd, Ctd, tud, tptd = (torch.from_numpy(v).to(cuda) for v in (w, b, Cs, Ct, tu, tpt))
    u0 = torch.empty(P, device=cuda)
    pt0 = torch.empty(P, device=cuda)
    W._ck(L.asets_nn_head(_p(xd), P, _p(wd), _p(bd), _p(Csd), _p(Ctd), _p(u0), _p(pt0), None))
    dx = torch.empty(P * 16, dtype=torch.bfloat16, device=cuda)
    grads = torch.empty(35, device=cuda)
    nb = L.asets_nn_head_scratch_bytes(P)
    scratch = torch.empty(nb, dtype=torch.uint8, device=cuda)
    W._ck(L.asets_nn_head_bwd(_p(xd), P, _p(wd), _p(u0), _p(pt0), _p(tud), _p(tptd), _p(dx),
                              _p(grads), _p(scratch), nb, None))
    torch.cuda.synchronize()
    X = torch.as_tensor(x.astype(np.float64)).requires_grad_(True)
    Wt = torch.as_tensor(w.astype(np.float64)).requires_grad_(True)
    Bt = torch.as_tensor(b.astype(np.float64)).requires_grad_(True)
    a_ = X @ Wt[:16] + Bt[0]
    b_ = X @ Wt[16:] + Bt[1]
    ru0 = a_
    rpt0 = torch.as_tensor(np.minimum(Cs, Ct), dtype=torch.float64) + NO.COST_SCALE * b_
    loss = ((ru0 - torch.as_tensor(tu, dtype=torch.float64)) ** 2).mean() + \
        (((rpt0 - torch.as_tensor(tpt, dtype=torch.float64)) / NO.COST_SCALE) ** 2).mean()
    loss.backward()
    np.testing.assert_allclose(u0.cpu().numpy(), ru0.detach().numpy(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(pt0.cpu().numpy(), rpt0.detach().numpy(), rtol=1e-5, atol=1e-4)
    g = grads.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(g[:32], Wt.grad.numpy(), rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(g[32:34], Bt.grad.numpy(), rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(g[34] / P, loss.item(), rtol=1e-5)
    _assert_bf16_close(_f(dx), X.grad.numpy().ravel())


def test_adamw_matches_formula(cuda):
    rng = np.random.default_rng(9)
    n = 3000
    w = rng.normal(0, 1, n).astype(np.float32)
    g = rng.normal(0, 1, n).astype(np.float32)
    wd_, gd = torch.from_numpy(w.copy()).to(cuda), torch.from_numpy(g).to(cuda)
    m = torch.zeros(n, device=cuda)
    v = torch.zeros(n, device=cuda)
    wb = torch.empty(n, dtype=torch.bfloat16, device=cuda)
    L = W.nnlib()
    ww, mm, vv = w.astype(np.float64), np.zeros(n), np.zeros(n)
    for step in (1, 2, 3):
        W._ck(L.asets_nn_adamw(_p(wd_), _p(wb), _p(gd), _p(m), _p(v), n, 1e-3, 0.9, 0.999, 1e-8,
                               0.01, step, None))
        mm = 0.9 * mm + 0.1 * g
        vv = 0.999 * vv + 0.001 * g.astype(np.float64) ** 2
        mh, vh = mm / (1 - 0.9 ** step), vv / (1 - 0.999 ** step)
        ww = ww * (1 - 1e-3 * 0.01) - 1e-3 * mh / (np.sqrt(vh) + 1e-8)
    torch.cuda.synchronize()
    np.testing.assert_allclose(wd_.cpu().numpy(), ww, rtol=1e-5, atol=1e-6)
    assert np.array_equal(_f(wb), NO.bf16_round(wd_.cpu().numpy()))


def _net_inputs(shape, n, seed):
    from asetsmatlabmaxflow_amd import workloads

    rng = np.random.default_rng(seed)
    Cs, Ct, A, TU, TPT = [], [], [], [], []
    for v in range(n):
        cs, ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=seed + v, count=3)
        for lst, arr in ((Cs, cs), (Ct, ct), (A, a)):
            lst.append(np.asarray(arr, np.float32).ravel(order="F"))
        TU.append(rng.uniform(0, 1, cs.size).astype(np.float32))
        TPT.append((np.minimum(Cs[-1], Ct[-1]) + rng.normal(0, 2, cs.size)).astype(np.float32))
    return [np.concatenate(x) for x in (Cs, Ct, A, TU, TPT)]


@pytest.mark.parametrize("shape,n", [((16, 24, 8), 2), ((64, 8, 8), 2)])
def test_network_forward_and_gradients(cuda, shape, n):
    """Whole U-Net: loss and every parameter gradient of the project's kernels against
    float64 autograd through the oracle (bf16 rounding of activations mirrored in the
    forward pass; the activation gradients are bf16 only on the GPU, hence the 3 %
    relative-L2 bar per parameter tensor).  The second shape's 64-voxel rows run the
    LDS-staged convolution at full resolution."""
    Cs, Ct, A, TU, TPT = _net_inputs(shape, n, 77)
    net = W.WarmStartNet(cuda, seed=3)
    # non-zero biases so their gradients and the bias path are exercised
    flat = net.w32.cpu().numpy()
    rng = np.random.default_rng(1)
    for name, _, cout in W.CONVS:
        o, k = net.layout[name + ".b"]
        flat[o:o + k] = rng.normal(0, 0.05, k)
    o, k = net.layout["head.w"]
    flat[o:o + k] = rng.normal(0, 0.2, k)
    net.load_flat(flat)
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A, TU, TPT)]
    u0, pt0 = net.forward(d[0], d[1], d[2], shape, n, train=True)
    loss = net.backward(d[0], d[1], u0, pt0, d[3], d[4], shape, n)
    torch.cuda.synchronize()
    w32 = net.w32.cpu().numpy()
    w16 = net.w16.float().cpu().numpy()
    params = {}
    for key, (o, k) in net.layout.items():
        src = w16 if (key.endswith(".w") and not key.startswith("head")) else w32
        params[key] = torch.as_tensor(src[o:o + k].astype(np.float64)).requires_grad_(True)
    ru0, rpt0, rloss = NO.network(params, Cs, Ct, A, n, shape, TU, TPT)
    rloss.backward()
    np.testing.assert_allclose(u0.cpu().numpy(), ru0.detach().numpy(), rtol=0, atol=2e-2)
    np.testing.assert_allclose(loss.item(), rloss.item(), rtol=1e-2)
    grad = net.grad.cpu().numpy().astype(np.float64)
    worst = 0.0
    for key, (o, k) in net.layout.items():
        ref = params[key].grad.numpy()
        got = grad[o:o + k]
        rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30)
        worst = max(worst, rel)
        assert rel < 3e-2, f"{key}: relative L2 error {rel:.3e}"
    print(f"network gradient check: worst relative L2 error {worst:.3e}")


def test_training_step_many_boxes(cuda):
    """A training step where every level has more weight-gradient boxes than workgroups
    (scratch sized from every layer): finite loss and gradients, and the same bits when
    repeated (deterministic reductions)."""
    shape, n = (128, 64, 64), 2
    Cs, Ct, A, TU, TPT = _net_inputs(shape, n, 21)
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A, TU, TPT)]
    net = W.WarmStartNet(cuda, seed=2)
    grads, losses = [], []
    for _ in range(2):
        u0, pt0 = net.forward(d[0], d[1], d[2], shape, n, train=True)
        losses.append(float(net.backward(d[0], d[1], u0, pt0, d[3], d[4], shape, n)))
        grads.append(net.grad.clone())
    torch.cuda.synchronize()
    assert np.isfinite(losses).all() and torch.isfinite(grads[0]).all()
    assert losses[0] == losses[1] and torch.equal(grads[0], grads[1])


def test_training_reduces_loss(cuda):
    shape, n = (24, 24, 16), 2
    Cs, Ct, A, TU, TPT = _net_inputs(shape, n, 99)
    TU = (Cs > Ct).astype(np.float32)   # a learnable target: the initial cut
    TPT = np.minimum(Cs, Ct).astype(np.float32)
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A, TU, TPT)]
    net = W.WarmStartNet(cuda, seed=0)
    losses = []
    for _ in range(60):
        u0, pt0 = net.forward(d[0], d[1], d[2], shape, n, train=True)
        losses.append(float(net.backward(d[0], d[1], u0, pt0, d[3], d[4], shape, n)))
        net.adamw_step(lr=3e-3)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.5 * losses[0]


# ---- space-to-depth network (WarmStartNetS2D) ------------------------------------
def test_features_s2d(cuda):
    rng = np.random.default_rng(31)
    n, shape = 2, (6, 4, 8)          # (ny, nx, nz)
    P = n * int(np.prod(shape))
    Cs, Ct = rng.uniform(0, 30, P).astype(np.float32), rng.uniform(0, 30, P).astype(np.float32)
    a = rng.uniform(0.2, 1.0, P).astype(np.float32)
    feat = torch.empty(P * 4, dtype=torch.bfloat16, device=cuda)
    dev = [torch.from_numpy(v).to(cuda) for v in (Cs, Ct, a)]
    ny, nx, nz = shape
    W._ck(W.nnlib().asets_nn_features_s2d(*(_p(t) for t in dev), n, nz, nx, ny, _p(feat), None))
    torch.cuda.synchronize()
    want = NO.features_s2d(Cs, Ct, a, n, shape)
    assert np.array_equal(_f(feat).reshape(-1, 32), want)
    # sub-voxel s = (dz*2 + dx)*2 + dy of coarse voxel (0, 1, 1, 2) is fine (2+dz, 2+dx, 4+dy)
    c = ((0 * (nz // 2) + 1) * (nx // 2) + 1) * (ny // 2) + 2
    for s in range(8):
        dz, dx, dy = s >> 2, (s >> 1) & 1, s & 1
        g = ((2 + dz) * nx + 2 + dx) * ny + 4 + dy
        assert want[c, s * 4 + 1] == NO.bf16_round(np.float32(Ct[g] * np.float32(0.1)))


def test_head_s2d_forward_backward(cuda):
    rng = np.random.default_rng(32)
    L = W.nnlib()
    n, shape = 2, (10, 6, 4)
    ny, nx, nz = shape
    P = n * ny * nx * nz
    P2 = P // 8
    x = NO.bf16_round(np.maximum(rng.normal(0, 1, (P2, 16)), 0))
    w = rng.normal(0, 0.3, 512).astype(np.float32)
    b = rng.normal(0, 0.2, 32).astype(np.float32)
    tg = [rng.uniform(0, 1, P).astype(np.float32)] + \
        [rng.normal(0, 0.5, P).astype(np.float32) for _ in range(3)]
    lam = 2.0
    xd = _bf(x, cuda)
    wd, bd = torch.from_numpy(w).to(cuda), torch.from_numpy(b).to(cuda)
    tgd = [torch.from_numpy(t).to(cuda) for t in tg]
    out = [torch.empty(P, device=cuda) for _ in range(4)]
    W._ck(L.asets_nn_head_s2d(_p(xd), n, nz, nx, ny, _p(wd), _p(bd), *(_p(t) for t in out),
                              None))
    dx = torch.empty(P2 * 16, dtype=torch.bfloat16, device=cuda)
    grads = torch.empty(545, device=cuda)
    nb = L.asets_nn_head_s2d_scratch_bytes(n, nz, nx, ny)
    scratch = torch.empty(nb, dtype=torch.uint8, device=cuda)
    W._ck(L.asets_nn_head_s2d_bwd(_p(xd), n, nz, nx, ny, _p(wd), *(_p(t) for t in out),
                                  *(_p(t) for t in tgd), lam, _p(dx), _p(grads), _p(scratch),
                                  nb, None))
    torch.cuda.synchronize()
    X = torch.as_tensor(x.astype(np.float64)).requires_grad_(True)
    Wt = torch.as_tensor(w.astype(np.float64)).requires_grad_(True)
    Bt = torch.as_tensor(b.astype(np.float64)).requires_grad_(True)
    ro = NO.head_s2d(X, Wt, Bt, n, shape)
    t64 = [torch.as_tensor(t.astype(np.float64)) for t in tg]
    loss = ((ro[0] - t64[0]) ** 2).mean() + lam * sum(((ro[q] - t64[q]) ** 2).mean()
                                                      for q in (1, 2, 3))
    loss.backward()
    for q in range(4):
        np.testing.assert_allclose(out[q].cpu().numpy(), ro[q].detach().numpy(), rtol=1e-5,
                                   atol=1e-5)
    g = grads.cpu().numpy().astype(np.float64)
    np.testing.assert_allclose(g[:512], Wt.grad.numpy(), rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(g[512:544], Bt.grad.numpy(), rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(g[544] / P, loss.item(), rtol=1e-5)
    _assert_bf16_close(_f(dx), X.grad.numpy().ravel())


@pytest.mark.parametrize("n,shape", [(1, (5, 4, 3)), (2, (7, 3, 4)), (1, (9, 1, 1))])
def test_warm_state(cuda, n, shape):
    """ps and pt paired to a predicted (u, faces): bit for bit against the float32
    restatement, u clamped to [0, 1], zero-flux faces zeroed; the sink is saturated where
    u >= 1/2 and the source stays within Cs elsewhere."""
    rng = np.random.default_rng(33)
    P = n * int(np.prod(shape))
    u0 = rng.uniform(-0.2, 1.2, P).astype(np.float32)
    fl = [rng.normal(0, 0.4, P).astype(np.float32) for _ in range(3)]
    Cs, Ct = rng.uniform(0, 3, P).astype(np.float32), rng.uniform(0, 3, P).astype(np.float32)
    d = [torch.from_numpy(a.copy()).to(cuda) for a in [u0] + fl + [Cs, Ct]]
    ps, pt = torch.empty(P, device=cuda), torch.empty(P, device=cuda)
    ny, nx, nz = shape
    W._ck(W.nnlib().asets_nn_warm_state(*(_p(t) for t in d), ny, nx, nz, n, _p(ps), _p(pt),
                                        None))
    torch.cuda.synchronize()
    want = NO.warm_state(u0, *fl, Cs, Ct, n, shape)
    got = [d[0], ps, pt, d[1], d[2], d[3]]
    for g_, w_ in zip(got, want):
        assert np.array_equal(g_.cpu().numpy(), w_)
    u, ps_, pt_, bx, by, bz = want
    assert u.min() >= 0 and u.max() <= 1
    lo = u < 0.5
    assert np.all(ps_[lo] <= Cs[lo] + 1e-6)
    assert np.all(pt_[~lo] == Ct[~lo])
    f = lambda a: a.reshape(n, shape[2], shape[1], shape[0])
    assert not f(bx)[:, :, 0, :].any() and not f(by)[:, :, :, 0].any()
    assert not f(bz)[:, 0, :, :].any()


def _net_state_inputs(shape, n, seed):
    Cs, Ct, A, TU, _ = _net_inputs(shape, n, seed)
    rng = np.random.default_rng(seed + 1)
    P = Cs.size
    TB = [rng.normal(0, 0.3, P).astype(np.float32) for _ in range(3)]
    return Cs, Ct, A, TU, TB


@pytest.mark.parametrize("shape,n", [((32, 16, 16), 2), ((128, 16, 16), 1)])
def test_network_s2d_forward_and_gradients(cuda, shape, n):
    """Space-to-depth U-Net: predicted state, loss and every parameter gradient against
    float64 autograd through the oracle (3 % relative-L2 bar per parameter tensor, as for
    the full-resolution network).  The second shape's level-0 rows are 64 voxels wide
    (LDS-staged convolutions)."""
    Cs, Ct, A, TU, TB = _net_state_inputs(shape, n, 78)
    net = W.WarmStartNetS2D(cuda, seed=3)
    flat = net.w32.cpu().numpy()
    rng = np.random.default_rng(2)
    for name, _, cout in W.CONVS_S2D:
        o, k = net.layout[name + ".b"]
        flat[o:o + k] = rng.normal(0, 0.05, k)
    o, k = net.layout["head.w"]
    flat[o:o + k] = rng.normal(0, 0.2, k)
    net.load_flat(flat)
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A, TU, *TB)]
    pred = net.forward(d[0], d[1], d[2], shape, n, train=True)
    tgt = dict(zip(("u", "bx", "by", "bz"), d[3:]))
    loss = net.backward_state(pred, tgt, shape, n)
    torch.cuda.synchronize()
    w32 = net.w32.cpu().numpy()
    w16 = net.w16.float().cpu().numpy()
    params = {}
    for key, (o, k) in net.layout.items():
        src = w16 if (key.endswith(".w") and not key.startswith("head")) else w32
        params[key] = torch.as_tensor(src[o:o + k].astype(np.float64)).requires_grad_(True)
    rout, rloss = NO.network_s2d(params, Cs, Ct, A, n, shape, [TU, *TB], W.FLOW_LOSS_WEIGHT)
    rloss.backward()
    for q, k in enumerate(("u", "bx", "by", "bz")):
        np.testing.assert_allclose(pred[k].cpu().numpy(), rout[q].detach().numpy(), rtol=0,
                                   atol=2e-2)
    np.testing.assert_allclose(loss.item(), rloss.item(), rtol=1e-2)
    grad = net.grad.cpu().numpy().astype(np.float64)
    for key, (o, k) in net.layout.items():
        ref = params[key].grad.numpy()
        rel = np.linalg.norm(grad[o:o + k] - ref) / max(np.linalg.norm(ref), 1e-30)
        assert rel < 3e-2, f"{key}: relative L2 error {rel:.3e}"


def test_network_s2d_inference_equals_training_forward(cuda):
    shape, n = (64, 32, 16), 1
    Cs, Ct, A, _, _ = _net_state_inputs(shape, n, 6)
    net = W.WarmStartNetS2D(cuda, seed=5)
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A)]
    p1 = {k: v.clone() for k, v in net.forward(d[0], d[1], d[2], shape, n, train=True).items()}
    p2 = net.forward(d[0], d[1], d[2], shape, n, train=False)
    torch.cuda.synchronize()
    assert all(torch.equal(p1[k], p2[k]) for k in p1)


def test_network_s2d_training_reduces_loss(cuda):
    shape, n = (32, 32, 16), 2
    Cs, Ct, A, _, _ = _net_state_inputs(shape, n, 98)
    TU = (Cs > Ct).astype(np.float32)    # learnable targets: the initial cut, flows 0
    d = [torch.from_numpy(x).to(cuda) for x in (Cs, Ct, A, TU)]
    z = torch.zeros_like(d[3])
    net = W.WarmStartNetS2D(cuda, seed=0)
    losses = []
    for _ in range(60):
        pred = net.forward(d[0], d[1], d[2], shape, n, train=True)
        losses.append(float(net.backward_state(pred, {"u": d[3], "bx": z, "by": z, "bz": z},
                                               shape, n)))
        net.adamw_step(lr=3e-3)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.5 * losses[0]
