"""Hierarchical max-flow (HMF) on MI355X: the label-tree classes of the reference.

Mirrors maxflow/asetsHMF2D.m and asetsHMF3D.m (full-flow variant).  A tree is built
bottom-up exactly as in MATLAB::

    leaves = [asetsHMF2D([], alpha, Ct_k) for Ct_k in costs]      # asetsHMF2D({},alpha,Ct)
    root = asetsHMF2D(leaves, 0)                                 # asetsHMF2D(children,alpha)
    root.MaxFullFlow(200, 0.1, 1 / len(leaves))                  # numIts, steps, cc
    u_k = leaves[k].u

Arrays are given and returned in the reference's shapes (D1, D2[, D3]); inputs may be
numpy arrays or torch tensors, results are torch CUDA tensors (column-major views of
the device buffers: ``.cpu().numpy()`` gives the MATLAB-shaped array).  ``px``, ``py``,
``pz`` have the reference's shapes (D1-1, D2, D3), (D1, D2-1, D3), (D1, D2, D3-1).
The solve is one call to asets_hmf_dev (include/asets.h); there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .solver import _KERNELS, _stream_handle

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


def _flat_dev(a, shape, device="cuda") -> "torch.Tensor":
    """Column-major flat float32 device copy of an array of the given shape."""
    if torch.is_tensor(a):
        t = a.to(device=device, dtype=torch.float32)
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
        return t.permute(*range(t.dim() - 1, -1, -1)).contiguous().reshape(-1)
    a = np.asarray(a, dtype=np.float32)
    if a.shape != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {a.shape}")
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(device)


def _matlab_view(flat: "torch.Tensor", shape) -> "torch.Tensor":
    """The (D1, D2[, D3]) view of a column-major flat buffer."""
    nd = len(shape)
    return flat.view(*shape[::-1]).permute(*range(nd - 1, -1, -1))


@dataclass
class TreeResult:
    """Per-node device results of solve_tree (flat column-major, N floats each)."""
    u: list
    pt: list
    div: list
    faces: list          # per node: ndim own-face fields (dim 1, 2[, 3])
    l: list              # per node: normalised shape directions, or None
    data_cost: np.ndarray
    regu_cost: np.ndarray
    seconds: float
    kernel_ms: float | None = None


def solve_tree(parent, Ct, alpha, shape, num_it: int, steps: float, cc: float, *, l=None,
               kernel: str = "auto", stream=None, timing: bool = False,
               device="cuda") -> TreeResult:
    """MaxFullFlow on a tree given by parent indices (asets_hmf_dev).

    parent[n] is the parent index of node n (-1: root); children in index order.
    Ct[n]: data cost of leaf n (array of ``shape``) or None; alpha[n]: number or array;
    l[n]: None or the ndim shape-constraint directions (lx, ly[, lz]).
    """
    shape = tuple(int(d) for d in shape)
    if len(shape) not in (2, 3):
        raise ValueError("shape must be (D1, D2) or (D1, D2, D3)")
    nd = len(shape)
    ny, nx = shape[0], shape[1]
    nz = shape[2] if nd == 3 else 1
    N = ny * nx * nz
    M = len(parent)
    keep = []   # device buffers referenced by the node table
    nodes = (nat.HmfNode * M)()
    res = TreeResult([], [], [], [], [], np.zeros(M), np.zeros(M), 0.0)
    for n in range(M):
        e = nodes[n]
        e.parent = int(parent[n])
        if Ct[n] is not None:
            t = _flat_dev(Ct[n], shape, device)
            keep.append(t)
            e.Ct = t.data_ptr()
        a = alpha[n]
        if np.ndim(a) == 0 and not (torch.is_tensor(a) and a.dim() > 0):
            e.alpha = float(a)
        else:
            t = _flat_dev(a, shape, device)
            keep.append(t)
            e.alpha_map = t.data_ptr()
        ln = None
        if l is not None and l[n] is not None:
            if len(l[n]) != nd:
                raise ValueError(f"node {n}: {nd} shape directions expected")
            ln = []
            for k in range(nd):
                t = _flat_dev(l[n][k], shape, device)
                keep.append(t)
                e.l[k] = t.data_ptr()
                o = torch.empty(N, dtype=torch.float32, device=device)
                ln.append(o)
                e.l_out[k] = o.data_ptr()
        out = [torch.empty(N, dtype=torch.float32, device=device) for _ in range(3 + nd)]
        e.u, e.pt, e.div = (o.data_ptr() for o in out[:3])
        for k in range(nd):
            e.faces[k] = out[3 + k].data_ptr()
        res.u.append(out[0])
        res.pt.append(out[1])
        res.div.append(out[2])
        res.faces.append(out[3:])
        res.l.append(ln)
    secs = C.c_double(0.0)
    kms = C.c_double(0.0) if timing else None
    opts = nat.make_opts(stream=_stream_handle(stream), kernel=_KERNELS[kernel], kernel_ms=kms)
    nat.check(nat.lib().asets_hmf_dev(nd, ny, nx, nz, nodes, M, int(num_it), float(steps),
                                      float(cc), C.byref(secs), C.byref(opts)))
    res.data_cost = np.array([nodes[n].data_cost for n in range(M)])
    res.regu_cost = np.array([nodes[n].regu_cost for n in range(M)])
    res.seconds = secs.value
    res.kernel_ms = kms.value if kms is not None else None
    del keep
    return res


class _HMF:
    """Node of a hierarchical max-flow label tree (asetsHMF2D.m:29-76)."""

    NDIM = 2

    def __init__(self, children, alpha, Ct=None):
        self.Name = ""
        self.C = list(children) if children is not None else []
        for c in self.C:
            if not isinstance(c, type(self)):
                raise TypeError(f"children must be {type(self).__name__} objects")
            c.P = self
        self.P = None
        self.alpha = alpha
        if not self.C:
            if Ct is None:
                raise ValueError("a leaf needs Ct")
            self.Ct = Ct
            self.D = tuple(int(d) for d in np.shape(Ct))
        else:
            self.Ct = None
            self.D = self.C[0].D
        if len(self.D) != self.NDIM:
            raise ValueError(f"{type(self).__name__} needs {self.NDIM}-D arrays, got {self.D}")
        self.lx = self.ly = None
        if self.NDIM == 3:
            self.lz = None
        self.u = self.pt = self.div = None
        self.px = self.py = None
        if self.NDIM == 3:
            self.pz = None
        self.dataCost = self.reguCost = None

    # -- the tree in breadth-first order (children in cell order) --
    def _nodes(self):
        order = [self]
        k = 0
        while k < len(order):
            order.extend(order[k].C)
            k += 1
        return order

    def _directions(self):
        names = ("lx", "ly", "lz")[: self.NDIM]
        given = [getattr(self, a) for a in names]
        have = [g is not None and np.size(g) > 0 for g in given]
        if not any(have):
            return None
        if not all(have):
            raise ValueError("a shape constraint needs " + ", ".join(names))
        return given

    def MaxFullFlow(self, numIts, steps, cc, *, kernel="auto", stream=None):
        """Full-flow HMF (asetsHMF2D.m:78-88): numIts iterations, then the costs."""
        if self.P is not None:
            raise NotImplementedError("MaxFullFlow runs on the root of the tree")
        order = self._nodes()
        index = {id(h): i for i, h in enumerate(order)}
        parent = [-1 if h.P is None else index[id(h.P)] for h in order]
        res = solve_tree(parent, [h.Ct for h in order], [h.alpha for h in order], self.D,
                         numIts, steps, cc, l=[h._directions() for h in order],
                         kernel=kernel, stream=stream)
        D = self.D
        for i, h in enumerate(order):
            h.u = _matlab_view(res.u[i], D)
            h.pt = _matlab_view(res.pt[i], D)
            h.div = _matlab_view(res.div[i], D)
            f = [_matlab_view(t, D) for t in res.faces[i]]
            h.px = f[0][1:]
            h.py = f[1][:, 1:]
            if self.NDIM == 3:
                h.pz = f[2][:, :, 1:]
            if res.l[i] is not None:
                for a, t in zip(("lx", "ly", "lz"), res.l[i]):
                    setattr(h, a, _matlab_view(t, D))
            h.dataCost = float(res.data_cost[i])
            h.reguCost = float(res.regu_cost[i])
        return res

    def MaxPseudoFlow(self, numIts, steps, cc):
        raise NotImplementedError("only the full-flow variant (MaxFullFlow) is provided")


class asetsHMF2D(_HMF):
    """asetsHMF2D(children, alpha, Ct): maxflow/asetsHMF2D.m."""
    NDIM = 2


class asetsHMF3D(_HMF):
    """asetsHMF3D(children, alpha, Ct): maxflow/asetsHMF3D.m."""
    NDIM = 3
