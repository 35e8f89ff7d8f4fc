"""ctypes binding of libasets_hip.so (include/asets.h).

The library is loaded from this package directory only; if it is missing the import of
any solver entry fails loudly -- there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libasets_hip.so"

ASETS_OK, ASETS_EBADARG, ASETS_EALPHA, ASETS_ENOMEM, ASETS_EDEVICE, ASETS_ETOOBIG = range(6)
FLOW_F32, FLOW_F16 = 0, 1
KERNEL_AUTO, KERNEL_FUSED, KERNEL_PASSES = 0, 1, 2

# every extern "C" symbol declared in include/asets.h
EXPORTED = (
    "asets_binary2d", "asets_binary3d", "asets_binary_warm", "asets_potts2d", "asets_potts3d",
    "asets_ishikawa2d", "asets_ishikawa3d", "asets_ishikawa_dev",
    "asets_binary_dev", "asets_binary_dev_state", "asets_binary_dev_resume",
    "asets_binary_dev_noalloc",
    "asets_potts_dev",
    "asets_binary3d_batched_dev", "asets_binary_energy_dev", "asets_potts_energy_dev",
    "asets_binary_bounds_dev",
    "asets_cost_loglik_dev", "asets_cost_neglog_dev", "asets_alpha_contrast_dev",
    "asets_labels_dev", "asets_cost_l1_dev", "asets_cost_smooth_labels_dev", "asets_hmf_dev",
    "asets_selftest_division", "asets_hbm_stream_dev", "asets_last_error",
    "asets_version", "asets_release",
)


class Params(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("err_bound", C.c_float), ("cc", C.c_float),
                ("steps", C.c_float)]


class Opts(C.Structure):
    _fields_ = [("stream", C.c_void_p), ("flow_dtype", C.c_int32), ("kernel", C.c_int32),
                ("poll_every", C.c_int32), ("tile_x", C.c_int32), ("zchunk", C.c_int32),
                ("reserved", C.c_int32 * 3), ("kernel_ms", C.POINTER(C.c_double))]


class Volume(C.Structure):
    _fields_ = [("Cs", C.c_void_p), ("Ct", C.c_void_p), ("alpha", C.c_void_p),
                ("ny", C.c_int64), ("nx", C.c_int64), ("nz", C.c_int64),
                ("u", C.c_void_p), ("cvg", C.c_void_p), ("it_num", C.c_void_p)]


class HmfNode(C.Structure):
    _fields_ = [("parent", C.c_int32), ("alpha", C.c_float), ("alpha_map", C.c_void_p),
                ("Ct", C.c_void_p), ("l", C.c_void_p * 3), ("u", C.c_void_p),
                ("pt", C.c_void_p), ("div", C.c_void_p), ("faces", C.c_void_p * 3),
                ("l_out", C.c_void_p * 3), ("data_cost", C.c_double),
                ("regu_cost", C.c_double)]


class AsetsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"asets status {status}: {msg}")
        self.status = status


_lib = None

P = C.c_void_p
I64 = C.c_int64
PP = C.POINTER(Params)
PO = C.POINTER(Opts)
PD = C.POINTER(C.c_double)
PI32 = C.POINTER(C.c_int32)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m asetsmatlabmaxflow_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        host_bin = [P, P, P]
        L.asets_binary2d.argtypes = host_bin + [I64, I64, PP, P, P, P, PD, PO]
        L.asets_binary3d.argtypes = host_bin + [I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_binary_warm.argtypes = [C.c_int] + host_bin + [I64, I64, I64, PP, P, P, P, P, P,
                                                               PD, PO]
        L.asets_potts2d.argtypes = [P, P, I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_potts3d.argtypes = [P, P, I64, I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_binary_dev.argtypes = [C.c_int, P, P, P, I64, I64, I64, PP, P, P, P, P, P, PD,
                                       PO]
        L.asets_binary_dev_noalloc.argtypes = [C.c_int, P, P, P, I64, I64, I64, PP, P, P, P,
                                               P, P, PD, PO]
        L.asets_binary_dev_state.argtypes = [C.c_int, P, P, P, I64, I64, I64, PP, P, P, P, P, P,
                                             PD, P, P, P, P, P, PO]
        L.asets_binary_dev_resume.argtypes = [C.c_int, P, P, P, I64, I64, I64, PP, P, P, P, P,
                                              P, P, P, P, P, PD, P, P, P, P, P, PO]
        L.asets_potts_dev.argtypes = [C.c_int, P, P, I64, I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_ishikawa2d.argtypes = [P, P, I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_ishikawa3d.argtypes = [P, P, I64, I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_ishikawa_dev.argtypes = [C.c_int, P, P, I64, I64, I64, I64, PP, P, P, P, PD, PO]
        L.asets_binary3d_batched_dev.argtypes = [C.c_int32, C.POINTER(Volume), PP, PD, PO]
        L.asets_binary_energy_dev.argtypes = [C.c_int, P, P, P, I64, I64, I64, P, PD, PO]
        L.asets_potts_energy_dev.argtypes = [C.c_int, P, P, I64, I64, I64, I64, P, PD, PO]
        L.asets_binary_bounds_dev.argtypes = [C.c_int, P, P, P, P, P, P, P, I64, I64, I64,
                                              C.c_int32, PD, PO]
        L.asets_cost_loglik_dev.argtypes = [P, P, I64, C.c_int32, C.c_double, P, PO]
        L.asets_cost_neglog_dev.argtypes = [P, I64, C.c_double, P, PO]
        L.asets_alpha_contrast_dev.argtypes = [C.c_int, P, I64, I64, I64, C.c_double, C.c_double,
                                               C.c_double, C.c_int, C.c_int, P, PO]
        L.asets_labels_dev.argtypes = [C.c_int, P, I64, C.c_int32, P, PO]
        L.asets_cost_l1_dev.argtypes = [P, I64, P, C.c_int32, P, PO]
        L.asets_cost_smooth_labels_dev.argtypes = [P, I64, I64, P, C.c_int32, P, C.c_int32, P,
                                                   PO]
        L.asets_hmf_dev.argtypes = [C.c_int, I64, I64, I64, C.POINTER(HmfNode), C.c_int32,
                                    C.c_int32, C.c_float, C.c_double, PD, PO]
        L.asets_selftest_division.argtypes = [C.c_int, C.c_float, C.POINTER(C.c_uint64)]
        L.asets_hbm_stream_dev.argtypes = [C.c_int, P, I64, C.c_int32, C.c_int32, C.c_int32,
                                           PD, C.POINTER(C.c_int64), PO]
        for name in EXPORTED:
            if name in ("asets_last_error", "asets_version"):
                getattr(L, name).restype = C.c_char_p
            elif name == "asets_release":
                getattr(L, name).restype = None
            else:
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != ASETS_OK:
        raise AsetsError(status, lib().asets_last_error().decode())


def make_params(max_iter: int, err_bound: float, cc: float, steps: float) -> Params:
    return Params(int(max_iter), float(err_bound), float(cc), float(steps))


def make_opts(stream: int | None = None, flow_dtype: int = FLOW_F32, kernel: int = KERNEL_AUTO,
              poll_every: int = 0, tile_x: int = 0, zchunk: int = 0,
              kernel_ms: C.c_double | None = None) -> Opts:
    o = Opts()
    o.stream = stream or None
    o.flow_dtype = flow_dtype
    o.kernel = kernel
    o.poll_every = poll_every
    o.tile_x = tile_x
    o.zchunk = zchunk
    o.kernel_ms = C.pointer(kernel_ms) if kernel_ms is not None else None
    return o
