"""CPU oracle of the primal-dual bounds of the binary continuous max-flow model (TEST
INFRASTRUCTURE ONLY: the checker for asets_binary_bounds_dev).  FP64 numpy.

Model (asetsBinaryMF3D_mex.c): max sum ps s.t. ps <= Cs, pt <= Ct, div p - ps + pt = 0 and
the per-voxel flow capacity 0.5 * (sum of its six faces' p^2) <= alpha^2 (the norm of
projStep, :321-331).  The reference has no bound of its own; its only energy terms are
HMF's (asetsHMF2D.m:184-203).

* Dual: for feasible flows D(p) = sum min(Cs, Ct + div p).  The flows are made feasible
  face by face with the smaller feasibility factor of the face's two voxels (p~).
* Primal: for u in [0, 1], P(u) = sum (1-u) Cs + u Ct + sigma_K(grad u), with sigma_K the
  support function of the capacity set.  By Lagrange duality over the capacity
  constraints, sigma_K(w) = min over lambda >= 0 of
      F(lambda) = sum_v lambda_v alpha_v^2 + sum_f w_f^2 / (2 (lambda_v1(f) + lambda_v2(f))),
  so F at any lambda bounds it from above; red-black coordinate descent on lambda (a
  per-voxel Newton solve) drives F down to sigma_K.  The starting lambda, the equal split
  of every face, gives the voxel-centred TV bound.

D(p~) <= min P <= P_lambda(u) brackets the optimum.  Arrays are (Nz, Nx, Ny) views of
the column-major fields; faces use the solver's own-face convention (bx[v] is the face
between v - e_x and v, 0 on the low boundary).  The GPU kernels follow the same
formulas, sums over the six faces in the same order (k = x, y, z; own face, then the
+axis neighbour's).
"""
import numpy as np

AX = (1, 2, 0)   # bx along x (axis 1), by along y (axis 2), bz along z (axis 0)


def grid(a, shape):
    """Column-major flat (or already shaped) array -> float64 (Nz, Nx, Ny)."""
    ny, nx, nz = shape
    a = np.asarray(a, np.float64)
    return (a.ravel(order="F") if a.ndim > 1 else a).reshape(nz, nx, ny)


def hi(f, ax):
    """Value at the +ax neighbour (0 past the volume)."""
    g = np.zeros_like(f)
    sl = [slice(None)] * 3
    src = list(sl)
    sl[ax] = slice(None, -1)
    src[ax] = slice(1, None)
    g[tuple(sl)] = f[tuple(src)]
    return g


def lo(f, ax):
    """Value at the -ax neighbour (0 outside)."""
    g = np.zeros_like(f)
    sl = [slice(None)] * 3
    src = list(sl)
    sl[ax] = slice(1, None)
    src[ax] = slice(None, -1)
    g[tuple(sl)] = f[tuple(src)]
    return g


def norm(f):
    """Per-voxel flow norm sqrt(0.5 * sum of its six faces^2) (projStep's n)."""
    return np.sqrt(0.5 * sum(f[k] ** 2 + hi(f[k], AX[k]) ** 2 for k in range(3)))


def div(f):
    return sum(hi(f[k], AX[k]) - f[k] for k in range(3))


def dual(Cs, Ct, al, faces):
    """D(p~): the faces scaled to the smaller feasibility factor of their two voxels."""
    n = norm(faces)
    s = np.where(n > al, al / np.where(n > 0, n, 1.0), 1.0)
    sc = [faces[k] * np.minimum(s, lo(s, AX[k])) for k in range(3)]
    return np.minimum(Cs, Ct + div(sc)).sum()


def wfaces(u):
    """Face differences u(v) - u(v - e_k) on the own faces (0 on the low boundary)."""
    w = [u - lo(u, AX[k]) for k in range(3)]
    for k in range(3):
        idx = [slice(None)] * 3
        idx[AX[k]] = 0
        w[k][tuple(idx)] = 0.0
    return w


def F(lam, al, w):
    tot = (al ** 2 * lam).sum()
    for k in range(3):
        c = 0.5 * w[k] ** 2
        L = lam + lo(lam, AX[k])
        tot += np.where(c > 0, c / np.where(L > 0, L, 1.0), 0.0).sum()
    return tot


def sweep(lam, al, w, color, newton=40):
    """Minimise F over the voxels of one checkerboard colour, neighbours fixed."""
    cs, ms = [], []
    for k in range(3):
        c = 0.5 * w[k] ** 2
        cs += [c, hi(c, AX[k])]
        ms += [lo(lam, AX[k]), hi(lam, AX[k])]
    Cm, Mm = np.stack(cs), np.stack(ms)
    a2 = al ** 2
    csum = Cm.sum(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        d0 = a2 - np.where(Cm > 0, Cm / np.where(Mm > 0, Mm, 0.0) ** 2, 0.0).sum(0)
    d0 = np.where(np.isnan(d0), -np.inf, d0)
    active = (csum > 0) & (d0 < 0)
    mmin = np.where(Cm > 0, Mm, np.inf).min(0)
    x = np.where(active, np.sqrt(csum) / al - np.where(np.isfinite(mmin), mmin, 0.0), 1.0)
    x = np.where(x > 0, x, 1e-3)
    for _ in range(newton):
        den = x + Mm
        den = np.where(den > 0, den, 1.0)
        t = np.where(Cm > 0, Cm / den ** 2, 0.0)
        g = a2 - t.sum(0)
        h = 2.0 * np.where(Cm > 0, t / den, 0.0).sum(0)
        xn = x - g / np.where(h > 0, h, 1.0)
        x = np.where(xn > 0, xn, x / 4)
    x = np.where(active, x, 0.0)
    zz, xx, yy = np.indices(lam.shape)
    return np.where(((zz + xx + yy) % 2) == color, x, lam)


def primal_bound(Cs, Ct, al, u, sweeps=30):
    """(P_lambda(u) after `sweeps` red-black sweeps, the voxel-centred bound, lambda, w)."""
    ub = np.clip(u, 0.0, 1.0)
    w = wfaces(ub)
    nrm = np.sqrt(sum(w[k] ** 2 + hi(w[k], AX[k]) ** 2 for k in range(3)))
    lam = nrm / (2.0 * np.sqrt(2.0) * al)   # equal split of every face: the voxel-centred TV
    data = ((1 - ub) * Cs + ub * Ct).sum()
    vox = data + F(lam, al, w)
    for i in range(sweeps):
        lam = sweep(lam, al, w, i % 2)
    return data + F(lam, al, w), vox, lam, w
