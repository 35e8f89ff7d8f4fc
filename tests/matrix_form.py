"""An independent FP64 restatement of one ASETS iteration in linear-algebra form.

Checker for the oracle (tests only).  Instead of loops over voxels with explicit
neighbour offsets and boundary tests, it builds the discrete gradient of each axis as a
sparse matrix from Kronecker products of 1-D difference matrices and writes the
iteration with matrix algebra:

  D_a        gradient along axis a: one row per voxel v (its low face), row = e_v - e_{v-s}
             for interior faces, zero row for the face at axis index 0
  b_a       <- b_a + steps * D_a gk                     (flow gradient step)
  |p|_v^2    = 0.5 * sum_a (|D_a|^T (b_a * b_a))_v       (each voxel's incident faces)
  g          = 1 / (n/alpha if n > alpha else 1)
  b_a       <- b_a * 0.5 * (|D_a| g)                     (rescale by the two voxels' mean)
  div        = -sum_a D_a^T b_a                          (divergence = minus adjoint)

followed by the pointwise source/sink/multiplier updates of the binary model
(Yuan-Bae-Tai, CVPR 2010) or of the Potts model (Yuan et al., ECCV 2010).  The stencil
offsets, signs and boundary rows are therefore fixed by the adjoint relation of the
gradient, not by retyping the oracle's loops.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _d1(n: int) -> sp.csr_matrix:
    """1-D backward difference with a zero first row (no face below index 0)."""
    if n == 1:
        return sp.csr_matrix((1, 1))
    m = sp.diags([np.r_[0.0, np.ones(n - 1)], -np.ones(n - 1)], [0, -1], shape=(n, n))
    return sp.csr_matrix(m)


def gradients(shape):
    """Sparse gradient matrices (D_y, D_x[, D_z]) for column-major volumes of `shape`
    (Ny fastest: flat index = z*Nx*Ny + x*Ny + y)."""
    if len(shape) == 2:
        ny, nx = shape
        Iy, Ix = sp.identity(ny), sp.identity(nx)
        return [sp.kron(Ix, _d1(ny), format="csr"), sp.kron(_d1(nx), Iy, format="csr")]
    ny, nx, nz = shape
    Iy, Ix, Iz = sp.identity(ny), sp.identity(nx), sp.identity(nz)
    Dy = sp.kron(Iz, sp.kron(Ix, _d1(ny)), format="csr")
    Dx = sp.kron(Iz, sp.kron(_d1(nx), Iy), format="csr")
    Dz = sp.kron(_d1(nz), sp.kron(Ix, Iy), format="csr")
    return [Dy, Dx, Dz]


class Flow:
    """Spatial flow field p = (b_a) with the projection/rescale step."""

    def __init__(self, shape):
        self.D = gradients(shape)
        self.absD = [abs(d) for d in self.D]
        self.N = int(np.prod(shape))

    def step_project(self, b, gk, alpha, steps):
        b = [ba + steps * (D @ gk) for ba, D in zip(b, self.D)]
        n2 = 0.5 * sum(A.T @ (ba * ba) for ba, A in zip(b, self.absD))
        n = np.sqrt(n2)
        g = 1.0 / np.where(n > alpha, n / alpha, 1.0)
        b = [ba * (0.5 * (A @ g)) for ba, A in zip(b, self.absD)]
        div = -sum(D.T @ ba for ba, D in zip(b, self.D))
        return b, div


def binary(Cs, Ct, alpha, shape, iters, cc, steps, u0=None, pt0=None):
    """FP64 binary model; returns (u, ps, pt, [b_a], cvg) after `iters` iterations."""
    f = lambda a: np.asarray(a, dtype=np.float64).ravel(order="F")
    Cs, Ct, alpha = f(Cs), f(Ct), f(alpha)
    F = Flow(shape)
    if u0 is None:
        u = (Cs >= Ct).astype(np.float64)
        ps = np.minimum(Cs, Ct)
        pt = ps.copy()
    else:
        u, pt = f(u0), f(pt0)
        ps = pt.copy()
    b = [np.zeros(F.N) for _ in F.D]
    div = np.zeros(F.N)
    cvg = []
    for _ in range(iters):
        gk = div - (ps - pt + u / cc)
        b, div = F.step_project(b, gk, alpha, steps)
        ps = np.minimum(pt - u / cc + div + 1.0 / cc, Cs)
        pt = np.minimum(ps + u / cc - div, Ct)
        e = cc * (pt + div - ps)
        u = u - e
        cvg.append(np.abs(e).mean())
    return u, ps, pt, b, np.array(cvg)


def potts(Ct, alpha, shape, nlab, iters, cc, steps):
    """FP64 Potts model; Ct, alpha of shape shape+(L,); returns (u [N*L], ps, cvg)."""
    N = int(np.prod(shape))
    Ct = np.asarray(Ct, dtype=np.float64).reshape(N, nlab, order="F")
    alpha = np.asarray(alpha, dtype=np.float64).reshape(N, nlab, order="F")
    F = Flow(shape)
    ps = Ct.min(axis=1)
    lab = Ct.argmin(axis=1)
    u = np.zeros((N, nlab))
    u[np.arange(N), lab] = 1.0
    pt = np.repeat(ps[:, None], nlab, axis=1)
    b = [[np.zeros(N) for _ in F.D] for _ in range(nlab)]
    div = np.zeros((N, nlab))
    cvg = []
    for _ in range(iters):
        for l in range(nlab):
            gk = div[:, l] - (ps - pt[:, l] + u[:, l] / cc)
            b[l], div[:, l] = F.step_project(b[l], gk, alpha[:, l], steps)
            pt[:, l] = np.minimum(ps + u[:, l] / cc - div[:, l], Ct[:, l])
        ft = div + pt
        ps = (ft - u / cc).sum(axis=1) / nlab + 1.0 / (cc * nlab)
        e = cc * (ft - ps[:, None])
        u = u - e
        cvg.append(np.abs(e).mean())
    return u.ravel(order="F"), ps, np.array(cvg)
