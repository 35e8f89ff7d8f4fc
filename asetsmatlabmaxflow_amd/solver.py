"""Device-memory solver API over the C ABI (torch tensors are used as device buffers only).

Volumes are flat float32 tensors holding the column-major (MATLAB) layout: index
g = z*Nx*Ny + x*Ny + y, Ny fastest; a Potts field stores label l at offset l*N.  Use
``to_device``/``from_device`` to move numpy arrays of shape (Ny, Nx[, Nz][, L]) in and out.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat

try:
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None


@dataclass
class Result:
    u: "torch.Tensor"
    cvg: "torch.Tensor"
    it_num: int
    seconds: float
    kernel_ms: float | None = None
    state: dict = field(default_factory=dict)

    @property
    def iterations(self) -> int:
        """Iterations executed: it_num + 1 if the stop test fired, else max_iter."""
        return self._executed

    _executed: int = 0


def to_device(a: np.ndarray, device="cuda") -> "torch.Tensor":
    flat = np.asarray(a, dtype=np.float32).ravel(order="F")
    return torch.from_numpy(np.ascontiguousarray(flat)).to(device)


def from_device(t: "torch.Tensor", shape) -> np.ndarray:
    return t.detach().cpu().numpy().reshape(shape, order="F")


def _ptr(t) -> int | None:
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 CUDA tensor")
    return t.data_ptr()


_KERNELS = {"auto": nat.KERNEL_AUTO, "fused": nat.KERNEL_FUSED, "passes": nat.KERNEL_PASSES}
_FLOWS = {"f32": nat.FLOW_F32, "f16": nat.FLOW_F16}


def _stream_handle(stream) -> int | None:
    if stream is None:
        return torch.cuda.current_stream().cuda_stream
    return stream.cuda_stream


def binary(Cs, Ct, alpha, shape, max_iter: int, err_bound: float, cc: float, steps: float, *,
           u0=None, pt0=None, kernel: str = "auto", flow: str = "f32", tile_x: int = 0,
           zchunk: int = 0, poll_every: int = 0, stream=None, want_state: bool = False,
           time_kernel: bool = False, out=None) -> Result:
    """Binary continuous max-flow (asetsBinaryMF2D/3D_mex semantics) on device tensors."""
    if len(shape) == 2:
        ndim, (ny, nx), nz = 2, shape, 1
    elif len(shape) == 3:
        ndim, (ny, nx, nz) = 3, shape
    else:
        raise ValueError("shape must be (Ny, Nx) or (Ny, Nx, Nz)")
    N = ny * nx * nz
    for t in (Cs, Ct, alpha):
        if t.numel() != N:
            raise ValueError(f"array has {t.numel()} elements, shape implies {N}")
    dev = Cs.device
    u = out if out is not None else torch.empty(N, dtype=torch.float32, device=dev)
    cvg = torch.empty(max(max_iter, 1), dtype=torch.float32, device=dev)
    itn = torch.empty(1, dtype=torch.int32, device=dev)
    secs = C.c_double(0.0)
    kms = C.c_double(0.0) if time_kernel else None
    opts = nat.make_opts(_stream_handle(stream), _FLOWS[flow], _KERNELS[kernel], poll_every,
                         tile_x, zchunk, kms)
    p = nat.make_params(max_iter, err_bound, cc, steps)
    L = nat.lib()
    if want_state:
        st = {k: torch.empty(N, dtype=torch.float32, device=dev)
              for k in ("ps", "pt", "bx", "by", "bz")}
        rc = L.asets_binary_dev_state(ndim, _ptr(Cs), _ptr(Ct), _ptr(alpha), ny, nx, nz,
                                      C.byref(p), _ptr(u0), _ptr(pt0), _ptr(u), _ptr(cvg),
                                      itn.data_ptr(), C.byref(secs), _ptr(st["ps"]),
                                      _ptr(st["pt"]), _ptr(st["bx"]), _ptr(st["by"]),
                                      _ptr(st["bz"]), C.byref(opts))
    else:
        st = {}
        rc = L.asets_binary_dev(ndim, _ptr(Cs), _ptr(Ct), _ptr(alpha), ny, nx, nz, C.byref(p),
                                _ptr(u0), _ptr(pt0), _ptr(u), _ptr(cvg), itn.data_ptr(),
                                C.byref(secs), C.byref(opts))
    nat.check(rc)
    it = int(itn.item())
    r = Result(u=u, cvg=cvg[:max_iter], it_num=it, seconds=secs.value,
               kernel_ms=kms.value if kms is not None else None, state=st)
    r._executed = it + 1 if it < max_iter else max_iter
    return r


def binary_resume(Cs, Ct, alpha, shape, state: dict, max_iter: int, err_bound: float,
                  cc: float, steps: float, *, kernel: str = "auto", flow: str = "f32",
                  tile_x: int = 0, zchunk: int = 0, poll_every: int = 0, stream=None,
                  want_state: bool = False, time_kernel: bool = False, out=None) -> Result:
    """Continue a binary solve from a full state {u, ps, pt, bx, by, bz} (device tensors,
    the convention of Result.state); k iterations then a resume of m equal k+m iterations.
    Without "ps" and "pt" the library pairs them to u and the faces (the fixed point of
    the source/sink update, as asets_nn_warm_state)."""
    ndim, ny, nx, nz = _ndim_dims(shape)
    N = ny * nx * nz
    dev = Cs.device
    u = out if out is not None else torch.empty(N, dtype=torch.float32, device=dev)
    cvg = torch.empty(max(max_iter, 1), dtype=torch.float32, device=dev)
    itn = torch.empty(1, dtype=torch.int32, device=dev)
    secs = C.c_double(0.0)
    kms = C.c_double(0.0) if time_kernel else None
    opts = nat.make_opts(_stream_handle(stream), _FLOWS[flow], _KERNELS[kernel], poll_every,
                         tile_x, zchunk, kms)
    p = nat.make_params(max_iter, err_bound, cc, steps)
    st = {k: torch.empty(N, dtype=torch.float32, device=dev)
          for k in ("ps", "pt", "bx", "by", "bz")} if want_state else {}
    rc = nat.lib().asets_binary_dev_resume(
        ndim, _ptr(Cs), _ptr(Ct), _ptr(alpha), ny, nx, nz, C.byref(p), _ptr(state["u"]),
        _ptr(state.get("ps")), _ptr(state.get("pt")), _ptr(state["bx"]), _ptr(state["by"]),
        _ptr(state.get("bz") if ndim == 3 else None), _ptr(u), _ptr(cvg), itn.data_ptr(),
        C.byref(secs), *(_ptr(st.get(k)) for k in ("ps", "pt", "bx", "by", "bz")),
        C.byref(opts))
    nat.check(rc)
    it = int(itn.item())
    r = Result(u=u, cvg=cvg[:max_iter], it_num=it, seconds=secs.value,
               kernel_ms=kms.value if kms is not None else None, state=st)
    r._executed = it + 1 if it < max_iter else max_iter
    return r


def potts(Ct, alpha, shape, nlab: int, max_iter: int, err_bound: float, cc: float,
          steps: float, *, kernel: str = "auto", tile_x: int = 0, zchunk: int = 0,
          poll_every: int = 0, stream=None, time_kernel: bool = False) -> Result:
    """Potts continuous max-flow (asetsPotts2D/3D_mex semantics) on device tensors."""
    if len(shape) == 2:
        ndim, (ny, nx), nz = 2, shape, 1
    elif len(shape) == 3:
        ndim, (ny, nx, nz) = 3, shape
    else:
        raise ValueError("shape must be (Ny, Nx) or (Ny, Nx, Nz)")
    NL = ny * nx * nz * nlab
    for t in (Ct, alpha):
        if t.numel() != NL:
            raise ValueError(f"array has {t.numel()} elements, shape implies {NL}")
    dev = Ct.device
    u = torch.empty(NL, dtype=torch.float32, device=dev)
    cvg = torch.empty(max(max_iter, 1), dtype=torch.float32, device=dev)
    itn = torch.empty(1, dtype=torch.int32, device=dev)
    secs = C.c_double(0.0)
    kms = C.c_double(0.0) if time_kernel else None
    opts = nat.make_opts(_stream_handle(stream), nat.FLOW_F32, _KERNELS[kernel], poll_every,
                         tile_x, zchunk, kms)
    p = nat.make_params(max_iter, err_bound, cc, steps)
    rc = nat.lib().asets_potts_dev(ndim, _ptr(Ct), _ptr(alpha), ny, nx, nz, nlab, C.byref(p),
                                   _ptr(u), _ptr(cvg), itn.data_ptr(), C.byref(secs),
                                   C.byref(opts))
    nat.check(rc)
    it = int(itn.item())
    r = Result(u=u, cvg=cvg[:max_iter], it_num=it, seconds=secs.value,
               kernel_ms=kms.value if kms is not None else None)
    r._executed = it + 1 if it < max_iter else max_iter
    return r


def ishikawa(Ct, alpha, shape, nlab: int, max_iter: int, err_bound: float, cc: float,
             steps: float, *, tile_x: int = 0, zchunk: int = 0, poll_every: int = 0,
             stream=None, time_kernel: bool = False) -> Result:
    """Ishikawa continuous max-flow (asetsIshikawa2D/3D_mex semantics) on device tensors:
    Ct holds nlab label fields, alpha nlab-1 (a longer alpha is cut); u has nlab-1."""
    ndim, ny, nx, nz = _ndim_dims(shape)
    N = ny * nx * nz
    if Ct.numel() != N * nlab:
        raise ValueError(f"Ct has {Ct.numel()} elements, shape implies {N * nlab}")
    if alpha.numel() < N * (nlab - 1):
        raise ValueError("alpha needs nlab-1 label fields")
    dev = Ct.device
    u = torch.empty(N * (nlab - 1), dtype=torch.float32, device=dev)
    cvg = torch.empty(max(max_iter, 1), dtype=torch.float32, device=dev)
    itn = torch.empty(1, dtype=torch.int32, device=dev)
    secs = C.c_double(0.0)
    kms = C.c_double(0.0) if time_kernel else None
    opts = nat.make_opts(_stream_handle(stream), nat.FLOW_F32, nat.KERNEL_FUSED, poll_every,
                         tile_x, zchunk, kms)
    p = nat.make_params(max_iter, err_bound, cc, steps)
    nat.check(nat.lib().asets_ishikawa_dev(ndim, _ptr(Ct), _ptr(alpha), ny, nx, nz, nlab,
                                           C.byref(p), _ptr(u), _ptr(cvg), itn.data_ptr(),
                                           C.byref(secs), C.byref(opts)))
    it = int(itn.item())
    r = Result(u=u, cvg=cvg[:max_iter], it_num=it, seconds=secs.value,
               kernel_ms=kms.value if kms is not None else None)
    r._executed = it + 1 if it < max_iter else max_iter
    return r


def binary_batched(volumes, max_iter: int, err_bound: float, cc: float, steps: float, *,
                   flow: str = "f32", tile_x: int = 0, zchunk: int = 0, stream=None,
                   time_kernel: bool = False):
    """Many binary 3D volumes of uneven size in one graph-captured launch per iteration.

    volumes: list of (Cs, Ct, alpha, (ny, nx, nz)) with flat float32 CUDA tensors.
    Returns (results, seconds, loop_ms): one Result per volume (u, cvg, it_num)."""
    nv = len(volumes)
    if nv == 0:
        raise ValueError("need at least one volume")
    sizes = []
    for i, (Cs, Ct, alpha, shape) in enumerate(volumes):
        ny, nx, nz = shape
        N = ny * nx * nz
        for t in (Cs, Ct, alpha):
            if t.numel() != N:
                raise ValueError(f"volume {i}: array has {t.numel()} elements, shape implies {N}")
        sizes.append(N)
    dev = volumes[0][0].device
    # one allocation per output kind: u of every volume back to back, a cvg row and an
    # it_num slot per volume (one read-back of all iteration counts at the end)
    starts = [0]
    for N in sizes:
        starts.append(starts[-1] + (N + 63) // 64 * 64)   # 256-byte aligned slices
    u_all = torch.empty(starts[-1], dtype=torch.float32, device=dev)
    cvg_all = torch.empty((nv, max(max_iter, 1)), dtype=torch.float32, device=dev)
    itn_all = torch.empty(nv, dtype=torch.int32, device=dev)
    vols = (nat.Volume * nv)()
    us = []
    for i, (Cs, Ct, alpha, shape) in enumerate(volumes):
        ny, nx, nz = shape
        u = u_all[starts[i]:starts[i] + sizes[i]]
        us.append(u)
        vols[i] = nat.Volume(_ptr(Cs), _ptr(Ct), _ptr(alpha), ny, nx, nz, u.data_ptr(),
                             cvg_all[i].data_ptr(), itn_all[i:i + 1].data_ptr())
    secs = C.c_double(0.0)
    kms = C.c_double(0.0) if time_kernel else None
    handle = stream.cuda_stream if stream is not None else None
    opts = nat.make_opts(handle, _FLOWS[flow], nat.KERNEL_FUSED, 0, tile_x, zchunk, kms)
    p = nat.make_params(max_iter, err_bound, cc, steps)
    nat.check(nat.lib().asets_binary3d_batched_dev(nv, vols, C.byref(p), C.byref(secs),
                                                   C.byref(opts)))
    out = []
    for i, it in enumerate(itn_all.cpu().tolist()):
        r = Result(u=us[i], cvg=cvg_all[i, :max_iter], it_num=it, seconds=secs.value)
        r._executed = it + 1 if it < max_iter else max_iter
        out.append(r)
    return out, secs.value, (kms.value if kms is not None else None)


def _ndim_dims(shape):
    if len(shape) == 2:
        return 2, shape[0], shape[1], 1
    if len(shape) == 3:
        return 3, shape[0], shape[1], shape[2]
    raise ValueError("shape must be (Ny, Nx) or (Ny, Nx, Nz)")


def energy_binary(Cs, Ct, alpha, u, shape, *, stream=None) -> np.ndarray:
    """FP64 energy {E, data, regularisation} of the binary labelling u (device tensors)."""
    ndim, ny, nx, nz = _ndim_dims(shape)
    out = np.zeros(3)
    opts = nat.make_opts(_stream_handle(stream), nat.FLOW_F32, nat.KERNEL_AUTO, 0, 0, 0, None)
    nat.check(nat.lib().asets_binary_energy_dev(
        ndim, _ptr(Cs), _ptr(Ct), _ptr(alpha), ny, nx, nz, _ptr(u),
        out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(opts)))
    return out


def binary_bounds(Cs, Ct, alpha, u, bx, by, bz, shape, *, sweeps: int = 30,
                  stream=None) -> dict:
    """FP64 bracket of the binary model's optimum for the state (u, bx, by, bz) on the
    device (asets_binary_bounds_dev): {"primal": P(u) with the regulariser bounded after
    `sweeps` red-black lambda sweeps, "primal_tv": the voxel-centred bound, "dual":
    D(p~), "width": (primal - dual) / |dual|}.  bz may be None in 2D."""
    ndim, ny, nx, nz = _ndim_dims(shape)
    out = np.zeros(3)
    opts = nat.make_opts(_stream_handle(stream), nat.FLOW_F32, nat.KERNEL_AUTO, 0, 0, 0, None)
    nat.check(nat.lib().asets_binary_bounds_dev(
        ndim, _ptr(Cs), _ptr(Ct), _ptr(alpha), _ptr(u), _ptr(bx), _ptr(by), _ptr(bz), ny, nx,
        nz, sweeps, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(opts)))
    return {"primal": float(out[0]), "primal_tv": float(out[1]), "dual": float(out[2]),
            "width": float((out[0] - out[2]) / abs(out[2])) if out[2] else float("inf")}


def energy_potts(Ct, alpha, u, shape, nlab: int, *, stream=None) -> np.ndarray:
    """FP64 energy {E, data, regularisation} of the Potts label fields u (device tensors)."""
    ndim, ny, nx, nz = _ndim_dims(shape)
    out = np.zeros(3)
    opts = nat.make_opts(_stream_handle(stream), nat.FLOW_F32, nat.KERNEL_AUTO, 0, 0, 0, None)
    nat.check(nat.lib().asets_potts_energy_dev(
        ndim, _ptr(Ct), _ptr(alpha), ny, nx, nz, nlab, _ptr(u),
        out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(opts)))
    return out


def hbm_stream(mode: str, buf, *, nontemporal: bool = True, one_in_flight: bool = False,
               grid: int = 0, iters: int = 10, stream=None) -> tuple[float, int]:
    """Time a plain stream over the device tensor ``buf`` (asets_hbm_stream_dev):
    mode "read", "copy", "mix" (the binary solver's 36 B read + 24 B written per voxel over
    its eleven arrays, without the stencil), "mix_packed" (the same bytes through three
    interleaved arrays), "mix_state" (inputs in three arrays, the state interleaved) or
    "mix_blk" (those five arrays with the fused binary kernel's layout and 4- and 8-byte
    lane accesses).  Returns (ms per launch, bytes read plus written per launch)."""
    m = {"read": 0, "copy": 1, "mix": 2, "mix_packed": 3, "mix_state": 4, "mix_blk": 5}[mode]
    if not (buf.is_cuda and buf.is_contiguous()):
        raise ValueError("expected a contiguous CUDA tensor")
    ms, traffic = C.c_double(), C.c_int64()
    opts = nat.make_opts(_stream_handle(stream), nat.FLOW_F32, nat.KERNEL_AUTO, 0, 0, 0, None)
    flags = (1 if nontemporal else 0) | (2 if one_in_flight else 0)
    nat.check(nat.lib().asets_hbm_stream_dev(
        m, buf.data_ptr(), buf.numel() * buf.element_size(), flags, grid, iters,
        C.byref(ms), C.byref(traffic), C.byref(opts)))
    return ms.value, traffic.value
