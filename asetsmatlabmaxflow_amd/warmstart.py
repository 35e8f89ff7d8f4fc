"""Learned warm start for the binary solver: a 4-level 3D U-Net (16/32/64/64 channels,
3x3x3 convolutions, bf16 weights and activations on the gfx950 matrix cores, fp32
accumulation) mapping (Cs, Ct, alpha) to a starting point (u0, pt0).

No reference counterpart: the reference starts every solve from the initial cut of
asetsBinaryMF3D_mex.c:205-224 (u = [Cs >= Ct], ps = pt = min(Cs, Ct)).  The prediction
replaces only that initialisation (asets_binary_warm / asets_binary_dev with u0, pt0: ps
starts at pt0, the flows at 0); the iteration and its stop rule are the plain solver's.

Every forward and backward operation is a kernel of libasets_nn.so (include/asets_nn.h);
torch tensors serve only as device buffers.  Layout: a batch of n column-major volumes
(Ny, Nx, Nz) is NDHWC with D = Nz, H = Nx, W = Ny; all three extents must be multiples of 8
(three 2x poolings).

Network (ReLU after every 3x3x3 convolution):
    enc0: 8 -> 16 -> 16            full resolution
    enc1: pool, 16 -> 32 -> 32     1/2
    enc2: pool, 32 -> 64 -> 64     1/4
    enc3: pool, 64 -> 64 -> 64     1/8
    dec2: [up(enc3), enc2] 128 -> 64
    dec1: [up(dec2), enc1]  96 -> 32
    dec0: [up(dec1), enc0]  48 -> 16
    head: 1x1x1, 16 -> (u0, (pt0 - min(Cs, Ct)) / 10)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

from . import _native as nat

PKG = Path(__file__).resolve().parent
NN_LIB = PKG / "libasets_nn.so"

# name, cin, cout
CONVS = (("enc0a", 8, 16), ("enc0b", 16, 16), ("enc1a", 16, 32), ("enc1b", 32, 32),
         ("enc2a", 32, 64), ("enc2b", 64, 64), ("enc3a", 64, 64), ("enc3b", 64, 64),
         ("dec2", 128, 64), ("dec1", 96, 32), ("dec0", 48, 16))
HEAD_W, HEAD_B = 32, 2

# every extern "C" symbol declared in include/asets_nn.h
NN_EXPORTED = (
    "asets_nn_conv3", "asets_nn_conv3_upcat", "asets_nn_flip_transpose", "asets_nn_relu_mask",
    "asets_nn_wgrad_scratch_bytes", "asets_nn_conv3_wgrad", "asets_nn_pool2",
    "asets_nn_pool2_bwd", "asets_nn_up2", "asets_nn_up2_bwd", "asets_nn_chan_copy",
    "asets_nn_features", "asets_nn_head", "asets_nn_head_scratch_bytes", "asets_nn_head_bwd",
    "asets_nn_adamw", "asets_nn_to_bf16", "asets_nn_last_error",
)

_nn = None


class NNError(RuntimeError):
    pass


def nnlib() -> C.CDLL:
    """libasets_nn.so; raises if it has not been built (there is no fallback)."""
    global _nn
    if _nn is None:
        if not NN_LIB.exists():
            raise NNError(f"{NN_LIB} is missing: run python -m asetsmatlabmaxflow_amd.build")
        L = C.CDLL(str(NN_LIB))
        P, I, I64, F = C.c_void_p, C.c_int, C.c_int64, C.c_float
        L.asets_nn_conv3.argtypes = [P, I, I, I, I, I, P, P, I, I, P, P]
        L.asets_nn_conv3_upcat.argtypes = [P, I, P, I, I, I, I, I, P, P, I, I, P, P]
        L.asets_nn_flip_transpose.argtypes = [P, P, I, I, P]
        L.asets_nn_relu_mask.argtypes = [P, P, I64, P]
        L.asets_nn_wgrad_scratch_bytes.argtypes = [I, I, I, I, I, I]
        L.asets_nn_wgrad_scratch_bytes.restype = C.c_size_t
        L.asets_nn_conv3_wgrad.argtypes = [P, I, I, I, I, I, P, I, P, P, I, P, C.c_size_t, P]
        L.asets_nn_pool2.argtypes = [P, I, I, I, I, I, P, P]
        L.asets_nn_pool2_bwd.argtypes = [P, I, I, I, I, I, P, I, P]
        L.asets_nn_up2.argtypes = [P, I, I, I, I, I, P, I, I, P]
        L.asets_nn_up2_bwd.argtypes = [P, I, I, I, I, I, I, I, P, P]
        L.asets_nn_chan_copy.argtypes = [P, I, I, P, I, I, I, I64, I, P]
        L.asets_nn_features.argtypes = [P, P, P, I64, P, P]
        L.asets_nn_head.argtypes = [P, I64, P, P, P, P, P, P, P]
        L.asets_nn_head_scratch_bytes.argtypes = [I64]
        L.asets_nn_head_scratch_bytes.restype = C.c_size_t
        L.asets_nn_head_bwd.argtypes = [P, I64, P, P, P, P, P, P, P, P, C.c_size_t, P]
        L.asets_nn_adamw.argtypes = [P, P, P, P, P, I64, F, F, F, F, F, I, P]
        L.asets_nn_to_bf16.argtypes = [P, P, I64, P]
        L.asets_nn_last_error.restype = C.c_char_p
        for name in NN_EXPORTED:
            f = getattr(L, name)
            if name not in ("asets_nn_wgrad_scratch_bytes", "asets_nn_head_scratch_bytes",
                            "asets_nn_last_error"):
                f.restype = C.c_int
        _nn = L
    return _nn


def _ck(rc: int) -> None:
    if rc != 0:
        raise NNError(f"asets_nn status {rc}: {nnlib().asets_nn_last_error().decode()}")


def _p(t) -> int | None:
    return None if t is None else t.data_ptr()


def _align(n: int, a: int = 64) -> int:
    return (n + a - 1) // a * a


class WarmStartNet:
    """Parameters, optimiser state and the forward/backward passes of the U-Net.

    All parameters live in one flat fp32 buffer (conv weights [27][cout][cin], then biases,
    then the head), mirrored in bf16 for the kernels; AdamW updates both in one launch."""

    def __init__(self, device="cuda", seed: int = 0):
        self.device = torch.device(device)
        self.layout = {}
        off = 0
        for name, cin, cout in CONVS:
            self.layout[name + ".w"] = (off, 27 * cout * cin)
            off = _align(off + 27 * cout * cin)
        for name, cin, cout in CONVS:
            self.layout[name + ".b"] = (off, cout)
            off = _align(off + cout)
        self.layout["head.w"] = (off, HEAD_W)
        off = _align(off + HEAD_W)
        self.layout["head.b"] = (off, HEAD_B)
        off = _align(off + HEAD_B)
        self.nparam = off
        self.w32 = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.grad = torch.zeros_like(self.w32)
        self.m = torch.zeros_like(self.w32)
        self.v = torch.zeros_like(self.w32)
        self.w16 = torch.zeros(off, dtype=torch.bfloat16, device=self.device)
        self.wflip = {name: torch.zeros(27 * cin * cout, dtype=torch.bfloat16, device=self.device)
                      for name, cin, cout in CONVS}
        self.step_count = 0
        self._bufs = {}
        self._scratch = None
        self.init_weights(seed)

    # -- parameters ------------------------------------------------------------------
    def view(self, key, which="w32"):
        off, n = self.layout[key]
        return getattr(self, which)[off:off + n]

    def init_weights(self, seed: int) -> None:
        """He-normal convolution weights, zero biases; head starts at u0 = 0.5, pt0 = min."""
        rng = np.random.default_rng(seed)
        host = np.zeros(self.nparam, dtype=np.float32)
        for name, cin, cout in CONVS:
            off, n = self.layout[name + ".w"]
            fan_in = 27 * (5 if name == "enc0a" else cin)   # 5 non-zero input features
            host[off:off + n] = rng.normal(0.0, np.sqrt(2.0 / fan_in), n)
        off, n = self.layout["head.w"]
        host[off:off + n] = rng.normal(0.0, 0.01, n)
        host[self.layout["head.b"][0]] = 0.5
        self.load_flat(host)

    def load_flat(self, host: np.ndarray) -> None:
        self.w32.copy_(torch.from_numpy(np.ascontiguousarray(host, dtype=np.float32)))
        _ck(nnlib().asets_nn_to_bf16(_p(self.w32), _p(self.w16), self.nparam, self._stream()))
        self.m.zero_()
        self.v.zero_()
        self.step_count = 0

    def state_dict(self) -> dict:
        host = self.w32.cpu().numpy()
        return {k: host[o:o + n].copy() for k, (o, n) in self.layout.items()}

    def save(self, path) -> None:
        np.savez(path, **self.state_dict())

    def save_checkpoint(self, path) -> None:
        """Weights, AdamW moments and step count, to continue a training run exactly."""
        np.savez(path, w32=self.w32.cpu().numpy(), m=self.m.cpu().numpy(),
                 v=self.v.cpu().numpy(), step=np.array([self.step_count], np.int64))

    def load_checkpoint(self, path) -> None:
        with np.load(path, allow_pickle=False) as z:
            w32, m, v = z["w32"], z["m"], z["v"]
            if w32.size != self.nparam or m.size != self.nparam or v.size != self.nparam:
                raise ValueError(f"checkpoint holds {w32.size} parameters, expected "
                                 f"{self.nparam}")
            self.load_flat(w32)
            self.m.copy_(torch.from_numpy(np.ascontiguousarray(m, np.float32)))
            self.v.copy_(torch.from_numpy(np.ascontiguousarray(v, np.float32)))
            self.step_count = int(z["step"][0])

    def load(self, path) -> None:
        with np.load(path, allow_pickle=False) as z:
            host = np.zeros(self.nparam, dtype=np.float32)
            for k, (o, n) in self.layout.items():
                a = z[k].astype(np.float32).ravel()
                if a.size != n:
                    raise ValueError(f"{k}: {a.size} values, expected {n}")
                host[o:o + n] = a
        self.load_flat(host)

    # -- buffers ---------------------------------------------------------------------
    @staticmethod
    def _stream():
        return torch.cuda.current_stream().cuda_stream

    def _geom(self, n, shape):
        ny, nx, nz = shape
        if ny % 8 or nx % 8 or nz % 8:
            raise ValueError("every extent must be a multiple of 8")
        # NDHWC geometry per level: (n, d, h, w)
        return [(n, nz >> l, nx >> l, ny >> l) for l in range(4)]

    def _buffers(self, n, shape, train: bool):
        key = (n, tuple(shape), train)
        if key in self._bufs:
            return self._bufs[key]
        self._bufs = {k: v for k, v in self._bufs.items() if k[2] != train}  # keep one per mode
        G = self._geom(n, shape)
        vox = [g[0] * g[1] * g[2] * g[3] for g in G]
        bf = torch.bfloat16
        E = lambda l, c: torch.empty(vox[l] * c, dtype=bf, device=self.device)
        b = {"x0": E(0, 8), "e0a": E(0, 16), "e0": E(0, 16), "p1": E(1, 16), "e1a": E(1, 32),
             "e1": E(1, 32), "p2": E(2, 32), "e2a": E(2, 64), "e2": E(2, 64), "p3": E(3, 64),
             "e3a": E(3, 64), "e3": E(3, 64), "d2": E(2, 64), "d1": E(1, 32), "d0": E(0, 16)}
        if train:
            # the backward pass needs the decoder inputs materialised (the weight
            # gradient reads them); inference gathers them inside the convolution
            b.update({"c2": E(2, 128), "c1": E(1, 96), "c0": E(0, 48)})
            b.update({"g0_16a": E(0, 16), "g0_16b": E(0, 16), "g0_48": E(0, 48),
                      "g1_32a": E(1, 32), "g1_32b": E(1, 32), "g1_96": E(1, 96),
                      "g1_16": E(1, 16), "g2_64a": E(2, 64), "g2_64b": E(2, 64),
                      "g2_128": E(2, 128), "g2_32": E(2, 32), "g3_64a": E(3, 64),
                      "g3_64b": E(3, 64), "g3_64c": E(3, 64)})
            L = nnlib()
            # every convolution's weight gradient (level = the digit in its name) and the head
            need = max([L.asets_nn_wgrad_scratch_bytes(*G[int(nm[3])], ci, co)
                        for nm, ci, co in CONVS] + [L.asets_nn_head_scratch_bytes(vox[0])])
            b["scratch"] = torch.empty(need, dtype=torch.uint8, device=self.device)
            b["hgrad"] = torch.empty(64, dtype=torch.float32, device=self.device)
        self._bufs[key] = (G, vox, b)
        return self._bufs[key]

    # -- forward ---------------------------------------------------------------------
    def _conv(self, name, src, g, dst, relu=True):
        cin, cout = next((ci, co) for nm, ci, co in CONVS if nm == name)
        _ck(nnlib().asets_nn_conv3(_p(src), *g, cin, _p(self.view(name + ".w", "w16")),
                                   _p(self.view(name + ".b")), cout, int(relu), _p(dst),
                                   self._stream()))

    def _dec(self, name, coarse, ca, skip, cat, g, dst, train):
        """Decoder convolution over [up2(coarse), skip]: materialised into `cat` when
        training (the backward pass reads it), gathered in the convolution otherwise."""
        cin, cout = next((ci, co) for nm, ci, co in CONVS if nm == name)
        L, s = nnlib(), self._stream()
        if train:
            _ck(L.asets_nn_up2(_p(coarse), *g, ca, _p(cat), cin, 0, s))
            _ck(L.asets_nn_chan_copy(_p(cat), cin, ca, _p(skip), cin - ca, 0, cin - ca,
                                     g[0] * g[1] * g[2] * g[3], 0, s))
            self._conv(name, cat, g, dst)
        else:
            _ck(L.asets_nn_conv3_upcat(_p(coarse), ca, _p(skip), *g, cin,
                                       _p(self.view(name + ".w", "w16")),
                                       _p(self.view(name + ".b")), cout, 1, _p(dst), s))

    def forward(self, Cs, Ct, alpha, shape, n=1, *, u0=None, pt0=None, train=False):
        """Predict (u0, pt0) for n volumes of `shape` stored back to back in the flat
        float32 device tensors Cs, Ct, alpha (column-major each).  Returns (u0, pt0)."""
        G, vox, b = self._buffers(n, shape, train)
        P = vox[0]
        for t in (Cs, Ct, alpha):
            if t.numel() != P or t.dtype != torch.float32 or not t.is_cuda:
                raise ValueError(f"expected {P} float32 device values")
        L, s = nnlib(), self._stream()
        _ck(L.asets_nn_features(_p(Cs), _p(Ct), _p(alpha), P, _p(b["x0"]), s))
        self._conv("enc0a", b["x0"], G[0], b["e0a"])
        self._conv("enc0b", b["e0a"], G[0], b["e0"])
        _ck(L.asets_nn_pool2(_p(b["e0"]), *G[0], 16, _p(b["p1"]), s))
        self._conv("enc1a", b["p1"], G[1], b["e1a"])
        self._conv("enc1b", b["e1a"], G[1], b["e1"])
        _ck(L.asets_nn_pool2(_p(b["e1"]), *G[1], 32, _p(b["p2"]), s))
        self._conv("enc2a", b["p2"], G[2], b["e2a"])
        self._conv("enc2b", b["e2a"], G[2], b["e2"])
        _ck(L.asets_nn_pool2(_p(b["e2"]), *G[2], 64, _p(b["p3"]), s))
        self._conv("enc3a", b["p3"], G[3], b["e3a"])
        self._conv("enc3b", b["e3a"], G[3], b["e3"])
        self._dec("dec2", b["e3"], 64, b["e2"], b.get("c2"), G[2], b["d2"], train)
        self._dec("dec1", b["d2"], 64, b["e1"], b.get("c1"), G[1], b["d1"], train)
        self._dec("dec0", b["d1"], 32, b["e0"], b.get("c0"), G[0], b["d0"], train)
        if u0 is None:
            u0 = torch.empty(P, dtype=torch.float32, device=self.device)
        if pt0 is None:
            pt0 = torch.empty(P, dtype=torch.float32, device=self.device)
        _ck(L.asets_nn_head(_p(b["d0"]), P, _p(self.view("head.w")), _p(self.view("head.b")),
                            _p(Cs), _p(Ct), _p(u0), _p(pt0), s))
        return u0, pt0

    # -- backward --------------------------------------------------------------------
    def _conv_bwd(self, name, src, g, gout, gin, first=False):
        """Weight/bias gradients of conv `name` (input src) from gout; input gradient into
        gin unless first."""
        cin, cout = next((ci, co) for nm, ci, co in CONVS if nm == name)
        L, s = nnlib(), self._stream()
        scratch = self._cur["scratch"]
        _ck(L.asets_nn_conv3_wgrad(_p(src), *g, cin, _p(gout), cout,
                                   _p(self.view(name + ".w", "grad")),
                                   _p(self.view(name + ".b", "grad")), 0, _p(scratch),
                                   scratch.numel(), s))
        if not first:
            wf = self.wflip[name]
            _ck(L.asets_nn_flip_transpose(_p(self.view(name + ".w", "w16")), _p(wf), cout, cin,
                                          s))
            _ck(L.asets_nn_conv3(_p(gout), *g, cout, _p(wf), None, cin, 0, _p(gin), s))

    def backward(self, Cs, Ct, u0, pt0, tu, tpt, shape, n=1) -> "torch.Tensor":
        """Gradients of mean((u0-tu)^2) + mean(((pt0-tpt)/10)^2) into self.grad after a
        forward(train=True) on the same inputs.  Returns the loss (device scalar)."""
        G, vox, b = self._buffers(n, shape, True)
        self._cur = b
        L, s = nnlib(), self._stream()
        P = vox[0]
        R = lambda g, y, nel: _ck(L.asets_nn_relu_mask(_p(g), _p(y), nel, s))
        hg = b["hgrad"]
        _ck(L.asets_nn_head_bwd(_p(b["d0"]), P, _p(self.view("head.w")), _p(u0), _p(pt0),
                                _p(tu), _p(tpt), _p(b["g0_16a"]), _p(hg), _p(b["scratch"]),
                                b["scratch"].numel(), s))
        self.view("head.w", "grad").copy_(hg[:32])
        self.view("head.b", "grad").copy_(hg[32:34])
        # decoder level 0: d0 = relu(conv(c0))
        R(b["g0_16a"], b["d0"], P * 16)
        self._conv_bwd("dec0", b["c0"], G[0], b["g0_16a"], b["g0_48"])
        _ck(L.asets_nn_up2_bwd(_p(b["g0_48"]), 48, 0, *G[0], 32, _p(b["g1_32a"]), s))   # d d1
        _ck(L.asets_nn_chan_copy(_p(b["g0_16b"]), 16, 0, _p(b["g0_48"]), 48, 32, 16, P, 0, s))
        # decoder level 1
        R(b["g1_32a"], b["d1"], vox[1] * 32)
        self._conv_bwd("dec1", b["c1"], G[1], b["g1_32a"], b["g1_96"])
        _ck(L.asets_nn_up2_bwd(_p(b["g1_96"]), 96, 0, *G[1], 64, _p(b["g2_64a"]), s))   # d d2
        _ck(L.asets_nn_chan_copy(_p(b["g1_32b"]), 32, 0, _p(b["g1_96"]), 96, 64, 32, vox[1], 0,
                                 s))
        # decoder level 2
        R(b["g2_64a"], b["d2"], vox[2] * 64)
        self._conv_bwd("dec2", b["c2"], G[2], b["g2_64a"], b["g2_128"])
        _ck(L.asets_nn_up2_bwd(_p(b["g2_128"]), 128, 0, *G[2], 64, _p(b["g3_64a"]), s))  # d e3
        _ck(L.asets_nn_chan_copy(_p(b["g2_64b"]), 64, 0, _p(b["g2_128"]), 128, 64, 64, vox[2],
                                 0, s))
        # encoder level 3
        R(b["g3_64a"], b["e3"], vox[3] * 64)
        self._conv_bwd("enc3b", b["e3a"], G[3], b["g3_64a"], b["g3_64b"])
        R(b["g3_64b"], b["e3a"], vox[3] * 64)
        self._conv_bwd("enc3a", b["p3"], G[3], b["g3_64b"], b["g3_64c"])
        _ck(L.asets_nn_pool2_bwd(_p(b["g3_64c"]), *G[2], 64, _p(b["g2_64b"]), 1, s))   # += d e2
        # encoder level 2
        R(b["g2_64b"], b["e2"], vox[2] * 64)
        self._conv_bwd("enc2b", b["e2a"], G[2], b["g2_64b"], b["g2_64a"])
        R(b["g2_64a"], b["e2a"], vox[2] * 64)
        self._conv_bwd("enc2a", b["p2"], G[2], b["g2_64a"], b["g2_32"])
        _ck(L.asets_nn_pool2_bwd(_p(b["g2_32"]), *G[1], 32, _p(b["g1_32b"]), 1, s))    # += d e1
        # encoder level 1
        R(b["g1_32b"], b["e1"], vox[1] * 32)
        self._conv_bwd("enc1b", b["e1a"], G[1], b["g1_32b"], b["g1_32a"])
        R(b["g1_32a"], b["e1a"], vox[1] * 32)
        self._conv_bwd("enc1a", b["p1"], G[1], b["g1_32a"], b["g1_16"])
        _ck(L.asets_nn_pool2_bwd(_p(b["g1_16"]), *G[0], 16, _p(b["g0_16b"]), 1, s))    # += d e0
        # encoder level 0
        R(b["g0_16b"], b["e0"], P * 16)
        self._conv_bwd("enc0b", b["e0a"], G[0], b["g0_16b"], b["g0_16a"])
        R(b["g0_16a"], b["e0a"], P * 16)
        self._conv_bwd("enc0a", b["x0"], G[0], b["g0_16a"], None, first=True)
        return hg[34] / P

    def adamw_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.01) -> None:
        self.step_count += 1
        _ck(nnlib().asets_nn_adamw(_p(self.w32), _p(self.w16), _p(self.grad), _p(self.m),
                                   _p(self.v), self.nparam, lr, b1, b2, eps, wd,
                                   self.step_count, self._stream()))


def warm_solve(net: WarmStartNet, Cs, Ct, alpha, shape, max_iter, err_bound, cc, steps,
               **kw):
    """Predict (u0, pt0) with the network, then run the plain solver from there."""
    from . import solver

    u0, pt0 = net.forward(Cs, Ct, alpha, shape, 1)
    return solver.binary(Cs, Ct, alpha, shape, max_iter, err_bound, cc, steps, u0=u0, pt0=pt0,
                         **kw)
