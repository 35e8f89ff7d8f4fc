"""ctypes wrapper of oracle/liboracle.so -- the CPU checker (test infrastructure only).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use this module.
Inputs are numpy arrays in MATLAB layout (shape (Ny, Nx[, Nz][, L]), any order; they are
flattened in Fortran order).  See oracle/asets_oracle.c for what each function follows
in the reference.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

ORC_OK, ORC_EBADARG, ORC_EALPHA, ORC_ENOMEM = range(4)


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle error {code}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE), "CC=gcc"], check=True)
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        I64 = C.c_int64
        for name, real in (("orc_binary_f32", C.c_float), ("orc_binary_f64", C.c_double)):
            f = getattr(L, name)
            f.argtypes = [P, P, P, I64, I64, I64, C.c_int, C.c_int, real, real, real, P, P,
                          C.c_int, P, P, P, P, P, P, P, P, P]
            f.restype = C.c_int
        for name, real in (("orc_potts_f32", C.c_float), ("orc_potts_f64", C.c_double)):
            f = getattr(L, name)
            f.argtypes = [P, P, I64, I64, I64, I64, C.c_int, C.c_int, real, real, real, C.c_int,
                          P, P, P, P, P, P, P, P, P]
            f.restype = C.c_int
        for name, real in (("orc_ishikawa_f32", C.c_float), ("orc_ishikawa_f64", C.c_double)):
            f = getattr(L, name)
            f.argtypes = [P, P, I64, I64, I64, I64, C.c_int, C.c_int, real, real, real, C.c_int,
                          P, P, P, P, P, P, P, P]
            f.restype = C.c_int
        for name, real in (("orc_hmf_f32", C.c_float), ("orc_hmf_f64", C.c_double)):
            f = getattr(L, name)
            f.argtypes = [I64, I64, I64, C.c_int, C.c_int, P, P, P, P, P, C.c_int, real,
                          C.c_double, P, P, P, P, P, P, P]
            f.restype = C.c_int
        L.orc_energy_binary.argtypes = [P, P, P, I64, I64, I64, C.c_int, P, P]
        L.orc_energy_binary.restype = C.c_int
        L.orc_energy_potts.argtypes = [P, P, I64, I64, I64, I64, C.c_int, P, P]
        L.orc_energy_potts.restype = C.c_int
        _lib = L
    return _lib


@dataclass
class OracleResult:
    u: np.ndarray
    cvg_ref: np.ndarray
    cvg64: np.ndarray
    it_num: int
    state: dict = field(default_factory=dict)

    @property
    def executed(self) -> int:
        return self.it_num + 1 if self.it_num < len(self.cvg_ref) else len(self.cvg_ref)


def _flat(a, dt):
    return np.ascontiguousarray(np.asarray(a, dtype=dt).ravel(order="F"))


def _p(a):
    return None if a is None else a.ctypes.data


def binary(Cs, Ct, alpha, shape, max_iter, err_bound, cc, steps, *, precision="f32",
           u0=None, pt0=None, stop_rule=0, want_state=False, is3d=None) -> OracleResult:
    """Binary model.  shape (Ny, Nx) uses the 2D operation order of asetsBinaryMF2D_mex.c,
    (Ny, Nx, Nz) the 3D order of asetsBinaryMF3D_mex.c (is3d may force the 3D forms on a
    2-tuple-with-Nz=1 volume)."""
    dt = np.float32 if precision == "f32" else np.float64
    if len(shape) == 2:
        ny, nx = shape
        nz = 1
        d3 = False if is3d is None else is3d
    else:
        ny, nx, nz = shape
        d3 = True if is3d is None else is3d
    N = ny * nx * nz
    Cs, Ct, alpha = (_flat(x, dt) for x in (Cs, Ct, alpha))
    u0 = None if u0 is None else _flat(u0, dt)
    pt0 = None if pt0 is None else _flat(pt0, dt)
    u = np.zeros(N, dtype=dt)
    cr = np.zeros(max(max_iter, 1), dtype=np.float32)
    c64 = np.zeros(max(max_iter, 1), dtype=np.float64)
    itn = C.c_int32(0)
    st = {k: np.zeros(N, dtype=dt) for k in ("ps", "pt", "bx", "by", "bz")} if want_state else {}
    f = lib().orc_binary_f32 if dt == np.float32 else lib().orc_binary_f64
    rc = f(Cs.ctypes.data, Ct.ctypes.data, alpha.ctypes.data, ny, nx, nz, int(d3),
           int(max_iter), err_bound, cc, steps, _p(u0), _p(pt0), stop_rule, u.ctypes.data,
           cr.ctypes.data, c64.ctypes.data, C.byref(itn), *(_p(st.get(k)) for k in
                                                          ("ps", "pt", "bx", "by", "bz")))
    if rc != ORC_OK:
        raise OracleError(rc)
    return OracleResult(u, cr[:max_iter], c64[:max_iter], itn.value, st)


def potts(Ct, alpha, shape, nlab, max_iter, err_bound, cc, steps, *, precision="f32",
          stop_rule=0, want_state=False, is3d=None) -> OracleResult:
    dt = np.float32 if precision == "f32" else np.float64
    if len(shape) == 2:
        ny, nx = shape
        nz = 1
        d3 = False if is3d is None else is3d
    else:
        ny, nx, nz = shape
        d3 = True if is3d is None else is3d
    N = ny * nx * nz
    NL = N * nlab
    Ct, alpha = (_flat(x, dt) for x in (Ct, alpha))
    u = np.zeros(NL, dtype=dt)
    cr = np.zeros(max(max_iter, 1), dtype=np.float32)
    c64 = np.zeros(max(max_iter, 1), dtype=np.float64)
    itn = C.c_int32(0)
    st = {}
    if want_state:
        st = {"ps": np.zeros(N, dtype=dt)}
        st.update({k: np.zeros(NL, dtype=dt) for k in ("pt", "bx", "by", "bz")})
    f = lib().orc_potts_f32 if dt == np.float32 else lib().orc_potts_f64
    rc = f(Ct.ctypes.data, alpha.ctypes.data, ny, nx, nz, nlab, int(d3), int(max_iter),
           err_bound, cc, steps, stop_rule, u.ctypes.data, cr.ctypes.data, c64.ctypes.data,
           C.byref(itn), *(_p(st.get(k)) for k in ("ps", "pt", "bx", "by", "bz")))
    if rc != ORC_OK:
        raise OracleError(rc)
    return OracleResult(u, cr[:max_iter], c64[:max_iter], itn.value, st)


def ishikawa(Ct, alpha, shape, nlab, max_iter, err_bound, cc, steps, *, precision="f32",
             stop_rule=0, want_state=False, is3d=None) -> OracleResult:
    """Ishikawa model: Ct has nlab label fields, alpha (at least) nlab-1; u has nlab-1.
    2D stops on cvg < errBound, 3D on <= (as the two reference files)."""
    dt = np.float32 if precision == "f32" else np.float64
    ny, nx, nz, d3 = _dims(shape, is3d)
    N = ny * nx * nz
    M = nlab - 1
    Ct = _flat(Ct, dt)
    alpha = _flat(alpha, dt)
    if alpha.size < N * M or Ct.size != N * nlab:
        raise ValueError("Ct needs nlab fields and alpha at least nlab-1")
    u = np.zeros(N * M, dtype=dt)
    cr = np.zeros(max(max_iter, 1), dtype=np.float32)
    c64 = np.zeros(max(max_iter, 1), dtype=np.float64)
    itn = C.c_int32(0)
    st = {}
    if want_state:
        st = {"pt": np.zeros(N * nlab, dtype=dt)}
        st.update({k: np.zeros(N * M, dtype=dt) for k in ("bx", "by", "bz")})
    f = lib().orc_ishikawa_f32 if dt == np.float32 else lib().orc_ishikawa_f64
    rc = f(Ct.ctypes.data, alpha.ctypes.data, ny, nx, nz, nlab, int(d3), int(max_iter),
           err_bound, cc, steps, stop_rule, u.ctypes.data, cr.ctypes.data, c64.ctypes.data,
           C.byref(itn), *(_p(st.get(k)) for k in ("pt", "bx", "by", "bz")))
    if rc != ORC_OK:
        raise OracleError(rc)
    return OracleResult(u, cr[:max_iter], c64[:max_iter], itn.value, st)


def _dims(shape, is3d):
    if len(shape) == 2:
        return shape[0], shape[1], 1, (False if is3d is None else is3d)
    return shape[0], shape[1], shape[2], (True if is3d is None else is3d)


def energy_binary(Cs, Ct, alpha, u, shape, *, is3d=None):
    """FP64 primal energy (total, data, regularisation) of u; see orc_energy_binary."""
    ny, nx, nz, d3 = _dims(shape, is3d)
    Cs, Ct, alpha, u = (_flat(x, np.float64) for x in (Cs, Ct, alpha, u))
    out = np.zeros(3)
    rc = lib().orc_energy_binary(Cs.ctypes.data, Ct.ctypes.data, alpha.ctypes.data, ny, nx,
                                 nz, int(d3), u.ctypes.data, out.ctypes.data)
    if rc != ORC_OK:
        raise OracleError(rc)
    return out


def energy_potts(Ct, alpha, u, shape, nlab, *, is3d=None):
    """FP64 Potts energy (total, data, regularisation) of the L label fields u."""
    ny, nx, nz, d3 = _dims(shape, is3d)
    Ct, alpha, u = (_flat(x, np.float64) for x in (Ct, alpha, u))
    out = np.zeros(3)
    rc = lib().orc_energy_potts(Ct.ctypes.data, alpha.ctypes.data, ny, nx, nz, nlab, int(d3),
                                u.ctypes.data, out.ctypes.data)
    if rc != ORC_OK:
        raise OracleError(rc)
    return out


@dataclass
class HMFResult:
    """Per-node state after MaxFullFlow; arrays are (nnodes, *shape) in MATLAB layout.
    faces[n, k] is node n's own-face flow along dim k+1 (the reference's px, py, pz
    are faces[n, 0][1:, ...], faces[n, 1][:, 1:, ...], faces[n, 2][:, :, 1:])."""
    u: np.ndarray
    pt: np.ndarray
    div: np.ndarray
    faces: np.ndarray
    l: np.ndarray
    data_cost: np.ndarray
    regu_cost: np.ndarray


def hmf(parent, Ct, alpha, shape, num_it, steps, cc, *, l=None, precision="f32"):
    """Hierarchical max-flow, full-flow variant (asetsHMF2D.m / asetsHMF3D.m MaxFullFlow).

    parent[n]: parent index of node n (-1 for the root); children in index order.
    Ct[n]: data cost of leaf n (None elsewhere).  alpha[n]: scalar or array.
    l[n]: None or (lx, ly[, lz]) shape-constraint directions (asetsHMF3D.m:217-238).
    """
    shape = tuple(int(d) for d in shape)
    is3d = len(shape) == 3
    ny, nx = shape[0], shape[1]
    nz = shape[2] if is3d else 1
    N = ny * nx * nz
    M = len(parent)
    dt = np.float32 if precision == "f32" else np.float64
    par = np.ascontiguousarray(np.asarray(parent, dtype=np.int32))
    ct = np.zeros((M, N), dt)
    al = np.zeros((M, N), dt)
    li = np.zeros((M, 3, N), dt)
    hl = np.zeros(M, np.int32)
    for n in range(M):
        if Ct[n] is not None:
            ct[n] = _flat(np.broadcast_to(np.asarray(Ct[n]), shape), dt)
        al[n] = _flat(np.broadcast_to(np.asarray(alpha[n], dtype=dt), shape), dt)
        if l is not None and l[n] is not None:
            hl[n] = 1
            for k, q in enumerate(l[n]):
                li[n, k] = _flat(np.broadcast_to(np.asarray(q), shape), dt)
    u = np.zeros((M, N), dt)
    pt = np.zeros((M, N), dt)
    dv = np.zeros((M, N), dt)
    fc = np.zeros((M, 3, N), dt)
    lo = np.zeros((M, 3, N), dt)
    dcost = np.zeros(M)
    rcost = np.zeros(M)
    f = lib().orc_hmf_f32 if precision == "f32" else lib().orc_hmf_f64
    rc = f(ny, nx, nz, int(is3d), M, par.ctypes.data, ct.ctypes.data, al.ctypes.data,
           li.ctypes.data, hl.ctypes.data, int(num_it), steps, cc, u.ctypes.data,
           pt.ctypes.data, dv.ctypes.data, fc.ctypes.data, lo.ctypes.data, dcost.ctypes.data,
           rcost.ctypes.data)
    if rc != ORC_OK:
        raise OracleError(rc)

    def sh(a):
        return a.reshape(a.shape[:-1] + shape[::-1]).transpose(
            tuple(range(a.ndim - 1)) + tuple(range(a.ndim - 1 + len(shape) - 1, a.ndim - 2, -1)))
    return HMFResult(sh(u), sh(pt), sh(dv), sh(fc), sh(lo), dcost, rcost)
