"""Drive the mex-ABI host files (mex/asets*_mex.so) from Python, as MATLAB would.

Each function mirrors the reference call of the same name
(maxflow/asetsBinaryMF2D_mex.c:36, asetsBinaryMF3D_mex.c:39, asetsPotts2D_mex.c:42,
asetsPotts3D_mex.c:44): float32 column-major inputs and a float32 ``pars`` vector in,
``(u, erriter, numIter, time)`` out.  Arguments go through the project's mx shim
(csrc/mex/shim), so the exported ``mexFunction`` symbols are exercised exactly as
MATLAB calls them; errors raised with ``mexErrMsgTxt`` surface as ``MexError``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _native

MEX_DIR = Path(__file__).resolve().parent / "mex"
mxSINGLE_CLASS, mxUINT16_CLASS = 7, 11


class MexError(RuntimeError):
    pass


_shim = None
_fns: dict[str, C.CDLL] = {}


def _load_shim() -> C.CDLL:
    global _shim
    if _shim is None:
        _native.lib()  # fail loudly if the HIP library is missing
        path = MEX_DIR / "libmexshim.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `python -m asetsmatlabmaxflow_amd.build`")
        s = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
        s.mxCreateNumericArray.restype = C.c_void_p
        s.mxCreateNumericArray.argtypes = [C.c_size_t, C.POINTER(C.c_size_t), C.c_int, C.c_int]
        s.mxGetData.restype = C.c_void_p
        s.mxGetData.argtypes = [C.c_void_p]
        s.mxGetNumberOfElements.restype = C.c_size_t
        s.mxGetNumberOfElements.argtypes = [C.c_void_p]
        s.mxGetNumberOfDimensions.restype = C.c_size_t
        s.mxGetNumberOfDimensions.argtypes = [C.c_void_p]
        s.mxGetDimensions.restype = C.POINTER(C.c_size_t)
        s.mxGetDimensions.argtypes = [C.c_void_p]
        s.mxGetClassID.restype = C.c_int
        s.mxGetClassID.argtypes = [C.c_void_p]
        s.mxDestroyArray.argtypes = [C.c_void_p]
        s.mexshim_call.restype = C.c_int
        s.mexshim_call.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int,
                                   C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
        s.mexshim_set_quiet.argtypes = [C.c_int]
        s.mexshim_live_arrays.restype = C.c_long
        s.mexshim_live_arrays.argtypes = []
        _shim = s
    return _shim


def live_arrays() -> int:
    """mxArrays allocated through the shim and not yet freed (a leak check)."""
    return int(_load_shim().mexshim_live_arrays())


def set_quiet(quiet: bool) -> None:
    """Silence the one-line report each solver prints through mexPrintf."""
    _load_shim().mexshim_set_quiet(1 if quiet else 0)


def _mex(name: str):
    if name not in _fns:
        _load_shim()
        path = MEX_DIR / f"{name}.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `python -m asetsmatlabmaxflow_amd.build`")
        _fns[name] = C.CDLL(str(path))
    return _fns[name].mexFunction


def _to_mx(a) -> int:
    """numpy array -> new mxArray (class from dtype: float32 -> single)."""
    s = _load_shim()
    a = np.asarray(a)
    if a.dtype == np.float32:
        cls = mxSINGLE_CLASS
    elif a.dtype == np.uint16:
        cls = mxUINT16_CLASS
    else:
        raise TypeError(f"unsupported dtype {a.dtype}: pass single(...) arrays")
    shape = a.shape if a.ndim >= 2 else ((a.size, 1) if a.ndim == 1 else (1, 1))
    dims = (C.c_size_t * len(shape))(*shape)
    m = s.mxCreateNumericArray(len(shape), dims, cls, 0)
    if not m:
        raise MemoryError("mxCreateNumericArray failed")
    buf = np.asfortranarray(a.reshape(shape, order="F"))
    C.memmove(s.mxGetData(m), buf.ctypes.data, buf.nbytes)
    return m


def _from_mx(m: int) -> np.ndarray:
    s = _load_shim()
    nd = s.mxGetNumberOfDimensions(m)
    dp = s.mxGetDimensions(m)
    shape = tuple(dp[i] for i in range(nd))
    cls = s.mxGetClassID(m)
    dt = {mxSINGLE_CLASS: np.float32, mxUINT16_CLASS: np.uint16}[cls]
    n = int(np.prod(shape))
    out = np.empty(n, dtype=dt)
    if n:
        C.memmove(out.ctypes.data, s.mxGetData(m), out.nbytes)
    return out.reshape(shape, order="F")


def call(name: str, *args, nlhs: int = 4):
    """Call mexFunction of mex/<name>.so with numpy arguments; returns nlhs arrays."""
    s = _load_shim()
    fn = _mex(name)
    prhs = (C.c_void_p * max(len(args), 1))()
    made = []
    try:
        for i, a in enumerate(args):
            prhs[i] = _to_mx(a)
            made.append(prhs[i])
        plhs = (C.c_void_p * max(nlhs, 1))()
        err = C.create_string_buffer(1024)
        rc = s.mexshim_call(C.cast(fn, C.c_void_p), nlhs, plhs, len(args), prhs, err, 1024)
        if rc != 0:
            raise MexError(err.value.decode())
        outs = []
        for i in range(nlhs):
            outs.append(_from_mx(plhs[i]))
            s.mxDestroyArray(plhs[i])
        if nlhs == 0 and plhs[0]:
            s.mxDestroyArray(plhs[0])   # MATLAB's ans
        return outs
    finally:
        for m in made:
            s.mxDestroyArray(m)


def _single(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float32)


def asetsBinaryMF2D_mex(Cs, Ct, alpha, pars, u0=None, pt0=None):
    """[u, erriter, numIter, time] = asetsBinaryMF2D_mex(single(Cs), single(Ct), single(alpha),
    single([Ny; Nx; maxIt; errBound; cc; steps]))  (asetsBinaryMF2D_mex.c:36-129)."""
    args = [_single(Cs), _single(Ct), _single(alpha), _single(pars)]
    if u0 is not None:
        args += [_single(u0), _single(pt0)]
    u, cvg, it, t = call("asetsBinaryMF2D_mex", *args)
    return u, cvg, it, t


def asetsBinaryMF3D_mex(Cs, Ct, alpha, pars, u0=None, pt0=None):
    """[u, erriter, numIter, time] = asetsBinaryMF3D_mex(single(Cs), single(Ct), single(alpha),
    single([Ny; Nx; Nz; maxIt; errBound; cc; steps]))  (asetsBinaryMF3D_mex.c:39-134)."""
    args = [_single(Cs), _single(Ct), _single(alpha), _single(pars)]
    if u0 is not None:
        args += [_single(u0), _single(pt0)]
    u, cvg, it, t = call("asetsBinaryMF3D_mex", *args)
    return u, cvg, it, t


def asetsPotts2D_mex(Ct, alpha, pars):
    """[u, erriter, numIter, time] = asetsPotts2D_mex(single(Ct), single(alpha),
    single([Ny; Nx; L; maxIt; errBound; cc; steps]))  (asetsPotts2D_mex.c:42-135)."""
    return tuple(call("asetsPotts2D_mex", _single(Ct), _single(alpha), _single(pars)))


def asetsPotts3D_mex(Ct, alpha, pars):
    """[u, erriter, numIter, time] = asetsPotts3D_mex(single(Ct), single(alpha),
    single([Ny; Nx; Nz; L; maxIt; errBound; cc; steps]))  (asetsPotts3D_mex.c:44-141)."""
    return tuple(call("asetsPotts3D_mex", _single(Ct), _single(alpha), _single(pars)))


def asetsIshikawa2D_mex(Ct, alpha, pars):
    """[u, erriter, numIter, time] = asetsIshikawa2D_mex(single(Ct), single(alpha),
    single([Ny; Nx; L; maxIt; errBound; cc; steps]))  (asetsIshikawa2D_mex.c:40-133);
    u has L-1 cut functions."""
    return tuple(call("asetsIshikawa2D_mex", _single(Ct), _single(alpha), _single(pars)))


def asetsIshikawa3D_mex(Ct, alpha, pars):
    """[u, erriter, numIter, time] = asetsIshikawa3D_mex(single(Ct), single(alpha),
    single([Ny; Nx; Nz; L; maxIt; errBound; cc; steps]))  (asetsIshikawa3D_mex.c:42-139)."""
    return tuple(call("asetsIshikawa3D_mex", _single(Ct), _single(alpha), _single(pars)))
