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
    "asets_nn_adamw", "asets_nn_to_bf16", "asets_nn_last_error", "asets_nn_features_s2d",
    "asets_nn_head_s2d", "asets_nn_head_s2d_scratch_bytes", "asets_nn_head_s2d_bwd",
    "asets_nn_warm_state",
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
        L.asets_nn_features_s2d.argtypes = [P, P, P, I, I, I, I, P, P]
        L.asets_nn_head_s2d.argtypes = [P, I, I, I, I, P, P, P, P, P, P, P]
        L.asets_nn_head_s2d_scratch_bytes.argtypes = [I, I, I, I]
        L.asets_nn_head_s2d_scratch_bytes.restype = C.c_size_t
        L.asets_nn_head_s2d_bwd.argtypes = [P, I, I, I, I, P, P, P, P, P, P, P, P, P, F, P, P, P,
                                            C.c_size_t, P]
        L.asets_nn_warm_state.argtypes = [P, P, P, P, P, P, I64, I64, I64, I, P, P, P]
        for name in NN_EXPORTED:
            f = getattr(L, name)
            if name not in ("asets_nn_wgrad_scratch_bytes", "asets_nn_head_scratch_bytes",
                            "asets_nn_head_s2d_scratch_bytes", "asets_nn_last_error"):
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

    CONVS = CONVS
    HEAD_W, HEAD_B = HEAD_W, HEAD_B
    NFEAT = 8           # input feature channels at level 0
    FAN_IN0 = 5         # of which non-zero
    SHIFT = 0           # level l runs at the input extents >> (l + SHIFT)

    def __init__(self, device="cuda", seed: int = 0):
        self.device = torch.device(device)
        self.layout = {}
        off = 0
        for name, cin, cout in self.CONVS:
            self.layout[name + ".w"] = (off, 27 * cout * cin)
            off = _align(off + 27 * cout * cin)
        for name, cin, cout in self.CONVS:
            self.layout[name + ".b"] = (off, cout)
            off = _align(off + cout)
        self.layout["head.w"] = (off, self.HEAD_W)
        off = _align(off + self.HEAD_W)
        self.layout["head.b"] = (off, self.HEAD_B)
        off = _align(off + self.HEAD_B)
        self.nparam = off
        self.w32 = torch.zeros(off, dtype=torch.float32, device=self.device)
        self.grad = torch.zeros_like(self.w32)
        self.m = torch.zeros_like(self.w32)
        self.v = torch.zeros_like(self.w32)
        self.w16 = torch.zeros(off, dtype=torch.bfloat16, device=self.device)
        self.wflip = {name: torch.zeros(27 * cin * cout, dtype=torch.bfloat16, device=self.device)
                      for name, cin, cout in self.CONVS}
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
        for name, cin, cout in self.CONVS:
            off, n = self.layout[name + ".w"]
            fan_in = 27 * (self.FAN_IN0 if name == "enc0a" else cin)
            host[off:off + n] = rng.normal(0.0, np.sqrt(2.0 / fan_in), n)
        off, n = self.layout["head.w"]
        host[off:off + n] = rng.normal(0.0, 0.01, n)
        self._init_head_bias(host, self.layout["head.b"][0])
        self.load_flat(host)

    def _init_head_bias(self, host, off) -> None:
        host[off] = 0.5   # u0 starts at 0.5

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
        m = 8 << self.SHIFT
        if ny % m or nx % m or nz % m:
            raise ValueError(f"every extent must be a multiple of {m}")
        # NDHWC geometry per level: (n, d, h, w)
        return [(n, nz >> (l + self.SHIFT), nx >> (l + self.SHIFT), ny >> (l + self.SHIFT))
                for l in range(4)]

    def _buffers(self, n, shape, train: bool):
        key = (n, tuple(shape), train)
        if key in self._bufs:
            return self._bufs[key]
        self._bufs = {k: v for k, v in self._bufs.items() if k[2] != train}  # keep one per mode
        G = self._geom(n, shape)
        vox = [g[0] * g[1] * g[2] * g[3] for g in G]
        bf = torch.bfloat16
        E = lambda l, c: torch.empty(vox[l] * c, dtype=bf, device=self.device)
        b = {"x0": E(0, self.NFEAT), "e0a": E(0, 16), "e0": E(0, 16), "p1": E(1, 16), "e1a": E(1, 32),
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
                        for nm, ci, co in self.CONVS] + [self._head_scratch(n, shape, vox)])
            b["scratch"] = torch.empty(need, dtype=torch.uint8, device=self.device)
            b["hgrad"] = torch.empty(_align(self.HEAD_W + self.HEAD_B + 1), dtype=torch.float32,
                                     device=self.device)
        self._bufs[key] = (G, vox, b)
        return self._bufs[key]

    def _head_scratch(self, n, shape, vox):
        return nnlib().asets_nn_head_scratch_bytes(vox[0])

    # -- forward ---------------------------------------------------------------------
    def _conv(self, name, src, g, dst, relu=True):
        cin, cout = next((ci, co) for nm, ci, co in self.CONVS if nm == name)
        _ck(nnlib().asets_nn_conv3(_p(src), *g, cin, _p(self.view(name + ".w", "w16")),
                                   _p(self.view(name + ".b")), cout, int(relu), _p(dst),
                                   self._stream()))

    def _dec(self, name, coarse, ca, skip, cat, g, dst, train):
        """Decoder convolution over [up2(coarse), skip]: materialised into `cat` when
        training (the backward pass reads it), gathered in the convolution otherwise."""
        cin, cout = next((ci, co) for nm, ci, co in self.CONVS if nm == name)
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

    def _trunk(self, b, G, train) -> None:
        """Encoder and decoder: b["x0"] -> b["d0"] (16 channels at level 0)."""
        L, s = nnlib(), self._stream()
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

    def _check_inputs(self, vox_fine, *ts):
        for t in ts:
            if t.numel() != vox_fine or t.dtype != torch.float32 or not t.is_cuda:
                raise ValueError(f"expected {vox_fine} float32 device values")

    def forward(self, Cs, Ct, alpha, shape, n=1, *, u0=None, pt0=None, train=False):
        """Predict (u0, pt0) for n volumes of `shape` stored back to back in the flat
        float32 device tensors Cs, Ct, alpha (column-major each).  Returns (u0, pt0)."""
        G, vox, b = self._buffers(n, shape, train)
        P = vox[0]
        self._check_inputs(P, Cs, Ct, alpha)
        L, s = nnlib(), self._stream()
        _ck(L.asets_nn_features(_p(Cs), _p(Ct), _p(alpha), P, _p(b["x0"]), s))
        self._trunk(b, G, train)
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
        cin, cout = next((ci, co) for nm, ci, co in self.CONVS if nm == name)
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
        hg = b["hgrad"]
        _ck(L.asets_nn_head_bwd(_p(b["d0"]), P, _p(self.view("head.w")), _p(u0), _p(pt0),
                                _p(tu), _p(tpt), _p(b["g0_16a"]), _p(hg), _p(b["scratch"]),
                                b["scratch"].numel(), s))
        self.view("head.w", "grad").copy_(hg[:32])
        self.view("head.b", "grad").copy_(hg[32:34])
        self._trunk_bwd(b, G, vox)
        return hg[34] / P

    def _trunk_bwd(self, b, G, vox) -> None:
        """Backward through decoder and encoder from d loss / d d0 in b["g0_16a"]."""
        L, s = nnlib(), self._stream()
        P = vox[0]
        R = lambda g, y, nel: _ck(L.asets_nn_relu_mask(_p(g), _p(y), nel, s))
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


# The space-to-depth network: the same trunk one level coarser (levels at 1/2 ... 1/16 of
# the input extents), fed the 2x2x2 blocks' costs as 32 channels and predicting the
# solver's whole state per voxel.
CONVS_S2D = (("enc0a", 32, 16),) + CONVS[1:]
FLOW_LOSS_WEIGHT = 2.0


class WarmStartNetS2D(WarmStartNet):
    """4-level U-Net (16/32/64/64 channels) on 2x2x2 voxel blocks: input features
    (Cs/10, Ct/10, alpha, (Ct-Cs)/10) of the block's 8 voxels (asets_nn_features_s2d),
    output (u0, bx0, by0, bz0) of the 8 voxels from a 1x1x1 head (asets_nn_head_s2d).
    With ps and pt paired to them (asets_nn_warm_state) that is a full solver state,
    resumed by asets_binary_dev_resume.  Every level runs at half the resolution of
    WarmStartNet's, so inference moves about an eighth of the activation bytes.
    Extents must be multiples of 16.

    A start from (u0, pt0) alone restarts the spatial flows at 0, and the flows are what
    the cold start spends its iterations building: from the exact converged (u, pt)
    with flows 0 the reference stop rule takes as many iterations as from the initial
    cut, while the exact full state stops after one."""

    CONVS = CONVS_S2D
    HEAD_W, HEAD_B = 32 * 16, 32
    NFEAT = 32
    FAN_IN0 = 32
    SHIFT = 1

    def _init_head_bias(self, host, off) -> None:
        host[off:off + 32:4] = 0.5   # u0 of the 8 sub-voxels starts at 0.5, flows at 0

    def _head_scratch(self, n, shape, vox):
        ny, nx, nz = shape
        return nnlib().asets_nn_head_s2d_scratch_bytes(n, nz, nx, ny)

    def forward(self, Cs, Ct, alpha, shape, n=1, *, out=None, train=False):
        """Predict (u0, bx0, by0, bz0) for n volumes of `shape` stored back to back in the
        flat float32 device tensors Cs, Ct, alpha.  Returns a dict of fine fields."""
        G, vox, b = self._buffers(n, shape, train)
        ny, nx, nz = shape
        P = n * ny * nx * nz
        self._check_inputs(P, Cs, Ct, alpha)
        L, s = nnlib(), self._stream()
        _ck(L.asets_nn_features_s2d(_p(Cs), _p(Ct), _p(alpha), n, nz, nx, ny, _p(b["x0"]), s))
        self._trunk(b, G, train)
        out = out or {}
        for k in ("u", "bx", "by", "bz"):
            if k not in out:
                out[k] = torch.empty(P, dtype=torch.float32, device=self.device)
        _ck(L.asets_nn_head_s2d(_p(b["d0"]), n, nz, nx, ny, _p(self.view("head.w")),
                                _p(self.view("head.b")), _p(out["u"]), _p(out["bx"]),
                                _p(out["by"]), _p(out["bz"]), s))
        return out

    def backward_state(self, pred: dict, target: dict, shape, n=1,
                       lam: float = FLOW_LOSS_WEIGHT) -> "torch.Tensor":
        """Gradients of (1/P) sum (u0-tu)^2 + lam*((bx0-tbx)^2 + (by0-tby)^2 + (bz0-tbz)^2)
        into self.grad after a forward(train=True) on the same inputs.  Returns the loss."""
        G, vox, b = self._buffers(n, shape, True)
        self._cur = b
        ny, nx, nz = shape
        L, s = nnlib(), self._stream()
        hg = b["hgrad"]
        _ck(L.asets_nn_head_s2d_bwd(
            _p(b["d0"]), n, nz, nx, ny, _p(self.view("head.w")), _p(pred["u"]), _p(pred["bx"]),
            _p(pred["by"]), _p(pred["bz"]), _p(target["u"]), _p(target["bx"]),
            _p(target["by"]), _p(target["bz"]), lam, _p(b["g0_16a"]), _p(hg),
            _p(b["scratch"]), b["scratch"].numel(), s))
        self.view("head.w", "grad").copy_(hg[:512])
        self.view("head.b", "grad").copy_(hg[512:544])
        self._trunk_bwd(b, G, vox)
        return hg[544] / (n * ny * nx * nz)

    def predict_state(self, Cs, Ct, alpha, shape, n=1) -> dict:
        """A full solver state {u, ps, pt, bx, by, bz} for asets_binary_dev_resume."""
        st = self.forward(Cs, Ct, alpha, shape, n)
        return pair_state(st, Cs, Ct, shape, n, stream=self._stream())


def pair_state(st: dict, Cs, Ct, shape, n=1, stream=None) -> dict:
    """Complete {u, bx, by, bz} (device tensors of n volumes, modified in place: u clamped
    to [0, 1], zero-flux faces set to 0) with the ps and pt that the reference's
    source/sink update pairs them with (asets_nn_warm_state): a full solver state."""
    ny, nx, nz = shape
    P = n * ny * nx * nz
    dev = st["u"].device
    st["ps"] = torch.empty(P, dtype=torch.float32, device=dev)
    st["pt"] = torch.empty(P, dtype=torch.float32, device=dev)
    _ck(nnlib().asets_nn_warm_state(_p(st["u"]), _p(st["bx"]), _p(st["by"]), _p(st["bz"]),
                                    _p(Cs), _p(Ct), ny, nx, nz, n, _p(st["ps"]),
                                    _p(st["pt"]),
                                    stream if stream is not None else WarmStartNet._stream()))
    return st


def warm_solve_state(net: WarmStartNetS2D, Cs, Ct, alpha, shape, max_iter, err_bound, cc,
                     steps, **kw):
    """Predict u and the face flows with the space-to-depth network, then resume the
    plain solver from them: asets_binary_dev_resume pairs ps and pt to them (as
    predict_state does) while it lays the state out."""
    from . import solver

    st = net.forward(Cs, Ct, alpha, shape, 1)
    return solver.binary_resume(Cs, Ct, alpha, shape, st, max_iter, err_bound, cc, steps, **kw)
