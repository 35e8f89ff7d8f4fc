"""The C-ABI library, the mx shim and the mex host files load and export what they
declare; argument validation that happens before any device call (CPU-only checks)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from asetsmatlabmaxflow_amd import _native as nat
from asetsmatlabmaxflow_amd import mex

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(asets_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = nat.lib()
    declared = _declared(ROOT / "include" / "asets.h")
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), f"libasets_hip.so does not export {name}"
    assert set(declared) == set(nat.EXPORTED)
    assert lib.asets_version().decode().startswith("asets-mi355x")


def test_nn_library_exports_every_declared_symbol():
    from asetsmatlabmaxflow_amd import warmstart

    lib = warmstart.nnlib()
    declared = _declared(ROOT / "include" / "asets_nn.h")
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), f"libasets_nn.so does not export {name}"
    assert set(declared) == set(warmstart.NN_EXPORTED)


def test_nn_bad_arguments_rejected_before_device_work():
    from asetsmatlabmaxflow_amd import warmstart

    L = warmstart.nnlib()
    # cin not a multiple of 8, unsupported cout, odd pool extent, scratch too small
    assert L.asets_nn_conv3(1, 1, 4, 4, 4, 12, 1, None, 16, 1, 1, None) == nat.ASETS_EBADARG
    assert L.asets_nn_conv3(1, 1, 4, 4, 4, 16, 1, None, 24, 1, 1, None) == nat.ASETS_EBADARG
    assert L.asets_nn_pool2(1, 1, 4, 5, 4, 16, 1, None) == nat.ASETS_EBADARG
    assert L.asets_nn_conv3_wgrad(1, 1, 4, 4, 4, 16, 1, 16, 1, None, 0, 1, 16,
                                  None) == nat.ASETS_EBADARG
    assert b"scratch" in L.asets_nn_last_error()
    assert L.asets_nn_conv3(1, 1 << 12, 1 << 12, 1 << 12, 8, 16, 1, None, 16, 1, 1,
                            None) == nat.ASETS_ETOOBIG
    # space-to-depth: odd extents, misaligned fields, too many elements
    assert L.asets_nn_features_s2d(8, 8, 8, 1, 4, 4, 5, 8, None) == nat.ASETS_EBADARG
    assert L.asets_nn_features_s2d(8, 12, 8, 1, 4, 4, 4, 8, None) == nat.ASETS_EBADARG
    assert L.asets_nn_head_s2d(8, 1, 4, 4, 4, 8, 8, 8, 8, 8, 8, None) == nat.ASETS_EBADARG
    assert L.asets_nn_features_s2d(8, 8, 8, 1, 1 << 10, 1 << 10, 1 << 10, 8,
                                   None) == nat.ASETS_ETOOBIG
    # warm state: every extent must fit the kernel's 32-bit coordinates
    one = 8
    assert L.asets_nn_warm_state(*([one] * 6), 1, 1, 1 << 31, 1, one, one,
                                 None) == nat.ASETS_EBADARG
    assert L.asets_nn_warm_state(*([one] * 6), 4, 4, 4, 0, one, one, None) == nat.ASETS_EBADARG
    assert L.asets_nn_warm_state(*([one] * 6), 4, 4, 4, 1, None, one, None) == nat.ASETS_EBADARG


def test_mex_modules_export_mexFunction():
    for name in mex.MEX_DIR.glob("asets*_mex.so"):
        assert hasattr(C.CDLL(str(name)), "mexFunction"), name
    assert len(list(mex.MEX_DIR.glob("asets*_mex.so"))) == 6
    shim = C.CDLL(str(mex.MEX_DIR / "libmexshim.so"))
    for sym in ("mxCreateNumericArray", "mxGetData", "mxGetNumberOfElements", "mexErrMsgTxt",
                "mexPrintf", "mexshim_call"):
        assert hasattr(shim, sym)


def _status(fn, *args):
    return fn(*args)


def test_bad_arguments_rejected_before_device_work():
    L = nat.lib()
    p = nat.make_params(10, 0.0, 0.25, 0.11)
    buf = np.zeros(64, np.float32)
    itn = C.c_int32(0)
    ptr = buf.ctypes.data
    # non-positive size
    assert L.asets_binary3d(ptr, ptr, ptr, 0, 4, 4, C.byref(p), ptr, ptr, C.byref(itn), None,
                            None) == nat.ASETS_EBADARG
    assert b"positive" in L.asets_last_error()
    # null array
    assert L.asets_binary3d(None, ptr, ptr, 4, 4, 4, C.byref(p), ptr, ptr, C.byref(itn), None,
                            None) == nat.ASETS_EBADARG
    # negative max_iter, zero cc
    bad = nat.make_params(-1, 0.0, 0.25, 0.11)
    assert L.asets_binary2d(ptr, ptr, ptr, 4, 4, C.byref(bad), ptr, ptr, C.byref(itn), None,
                            None) == nat.ASETS_EBADARG
    bad = nat.make_params(5, 0.0, 0.0, 0.11)
    assert L.asets_potts3d(ptr, ptr, 2, 2, 2, 2, C.byref(bad), ptr, ptr, C.byref(itn), None,
                           None) == nat.ASETS_EBADARG
    assert L.asets_binary_warm(2, ptr, ptr, ptr, 4, 4, 2, C.byref(p), None, None, ptr, ptr,
                               C.byref(itn), None, None) == nat.ASETS_EBADARG
    assert L.asets_binary_warm(3, ptr, ptr, ptr, 4, 4, 2, C.byref(p), ptr, None, ptr, ptr,
                               C.byref(itn), None, None) == nat.ASETS_EBADARG
    # resume: ps0 without pt0 (both or neither: neither pairs them), or no faces
    cvg = ptr
    assert L.asets_binary_dev_resume(3, ptr, ptr, ptr, 4, 4, 2, C.byref(p), ptr, ptr, None,
                                     ptr, ptr, ptr, ptr, cvg, ptr, None, None, None, None,
                                     None, None, None) == nat.ASETS_EBADARG
    assert b"both, or neither" in L.asets_last_error()
    assert L.asets_binary_dev_resume(3, ptr, ptr, ptr, 4, 4, 2, C.byref(p), ptr, None, None,
                                     ptr, None, ptr, ptr, cvg, ptr, None, None, None, None,
                                     None, None, None) == nat.ASETS_EBADARG
    # oversized plane
    assert L.asets_binary3d(ptr, ptr, ptr, 1 << 16, 1 << 15, 1, C.byref(p), ptr, ptr,
                            C.byref(itn), None, None) == nat.ASETS_ETOOBIG
    # bounds: bad ndim, 2D with nz > 1, missing bz in 3D, negative sweeps
    out = np.zeros(3)
    o = out.ctypes.data_as(C.POINTER(C.c_double))
    assert L.asets_binary_bounds_dev(4, ptr, ptr, ptr, ptr, ptr, ptr, ptr, 4, 4, 4, 3, o,
                                     None) == nat.ASETS_EBADARG
    assert L.asets_binary_bounds_dev(2, ptr, ptr, ptr, ptr, ptr, ptr, None, 4, 4, 2, 3, o,
                                     None) == nat.ASETS_EBADARG
    assert L.asets_binary_bounds_dev(3, ptr, ptr, ptr, ptr, ptr, ptr, None, 4, 4, 2, 3, o,
                                     None) == nat.ASETS_EBADARG
    assert L.asets_binary_bounds_dev(3, ptr, ptr, ptr, ptr, ptr, ptr, ptr, 4, 4, 2, -1, o,
                                     None) == nat.ASETS_EBADARG
    # batched: 1 to 65,535 volumes
    vols = (nat.Volume * 1)()
    assert L.asets_binary3d_batched_dev(0, vols, C.byref(p), None, None) == nat.ASETS_EBADARG
    assert L.asets_binary3d_batched_dev(65536, vols, C.byref(p), None,
                                        None) == nat.ASETS_EBADARG
    assert b"65535" in L.asets_last_error()


def test_mex_argument_validation():
    mex.set_quiet(True)
    f = np.zeros((4, 3, 2), np.float32)
    pars = np.array([4, 3, 2, 10, 0.0, 0.25, 0.11], np.float32)
    with pytest.raises(mex.MexError, match="expected"):
        mex.call("asetsBinaryMF3D_mex", f, f, f)
    with pytest.raises(mex.MexError, match="Cs has 24 elements"):
        mex.asetsBinaryMF3D_mex(f, f, f, np.array([5, 3, 2, 10, 0, 0.25, 0.11], np.float32))
    with pytest.raises(mex.MexError, match="single"):
        mex.call("asetsBinaryMF3D_mex", f, f, f, np.array([4, 3, 2, 10, 0, 0, 0], np.uint16))
    with pytest.raises(mex.MexError, match="pars needs 7"):
        mex.asetsBinaryMF3D_mex(f, f, f, pars[:5])
    neg = pars.copy()
    neg[3] = -1
    with pytest.raises(mex.MexError, match="maxIt"):
        mex.asetsBinaryMF3D_mex(f, f, f, neg)
    with pytest.raises(mex.MexError, match="Ny"):
        mex.asetsBinaryMF2D_mex(f, f, f, np.array([0, 3, 10, 0, 0.2, 0.16], np.float32))
    with pytest.raises(mex.MexError, match="expected"):
        mex.call("asetsPotts3D_mex", f, f)
    with pytest.raises(mex.MexError, match="Ct has 24"):
        mex.asetsPotts3D_mex(f, f, np.array([4, 3, 2, 2, 10, 0, 0.25, 0.11], np.float32))
    with pytest.raises(mex.MexError, match="u0"):
        mex.asetsBinaryMF3D_mex(f, f, f, pars, u0=f[:2], pt0=f)


def test_mex_temporaries_freed():
    """With nlhs below 2 erriter is a temporary of the mex function, and an error after
    the outputs exist (here: no device, so the solve fails) frees them: no mxArray is
    left behind either way, as MATLAB frees a MEX function's temporaries."""
    mex.set_quiet(True)
    f = np.ones((4, 3, 2), np.float32)
    pars = np.array([4, 3, 2, 3, 0.0, 0.25, 0.11], np.float32)
    base = mex.live_arrays()
    for nlhs in (0, 1, 2, 4):
        try:
            mex.call("asetsBinaryMF3D_mex", f, f, f, pars, nlhs=nlhs)
        except mex.MexError:
            pass
        assert mex.live_arrays() == base, f"nlhs={nlhs}"
    with pytest.raises(mex.MexError):
        mex.call("asetsBinaryMF3D_mex", f, f, f, pars[:3], nlhs=1)
    assert mex.live_arrays() == base


def test_mx_roundtrip():
    a = np.asfortranarray(np.arange(24, dtype=np.float32).reshape(4, 3, 2))
    m = mex._to_mx(a)
    back = mex._from_mx(m)
    mex._load_shim().mxDestroyArray(m)
    assert back.shape == (4, 3, 2) and np.array_equal(back, a)


def test_product_does_not_reference_the_oracle():
    """The product path (package + HIP sources) never imports, links or loads oracle/.
    (build.py only compiles the checker next to the product.)"""
    pkg = ROOT / "asetsmatlabmaxflow_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.c")) + \
            list(pkg.rglob("*.h")):
        if f.name == "build.py":
            continue
        text = f.read_text()
        assert "import oracle" not in text and "liboracle" not in text, f


def test_fused_size_limit_rejected_before_device_work():
    """The fused kernels need every field (plus two planes) below 2^31 bytes; an explicit
    fused request beyond that is refused before any device call (the per-pass kernels
    take over under ASETS_KERNEL_AUTO)."""
    L = nat.lib()
    p = nat.make_params(10, 0.0, 0.35, 0.11)
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    opts = nat.make_opts(None, nat.FLOW_F32, nat.KERNEL_FUSED, 0, 0, 0, None)
    rc = L.asets_binary_dev(3, ptr, ptr, ptr, 1024, 1024, 1024, C.byref(p), None, None, ptr,
                            ptr, ptr, None, C.byref(opts))
    assert rc == nat.ASETS_ETOOBIG and b"fused" in L.asets_last_error()
    opts16 = nat.make_opts(None, nat.FLOW_F16, nat.KERNEL_AUTO, 0, 0, 0, None)
    rc = L.asets_binary_dev(3, ptr, ptr, ptr, 1024, 1024, 1024, C.byref(p), None, None, ptr,
                            ptr, ptr, None, C.byref(opts16))
    assert rc == nat.ASETS_ETOOBIG
    rc = L.asets_ishikawa_dev(3, ptr, ptr, 1024, 1024, 1024, 3, C.byref(p), ptr, ptr, ptr, None,
                              None)
    assert rc == nat.ASETS_ETOOBIG


def test_hmf_tree_validation_before_device_work():
    L = nat.lib()
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data

    def call(parent, ct=None, l=None, ndim=2, nz=1, its=3, cc=0.3, kernel=nat.KERNEL_AUTO):
        M = len(parent)
        nodes = (nat.HmfNode * M)()
        for n, p in enumerate(parent):
            nodes[n].parent = p
            nodes[n].alpha = 0.1
            nodes[n].Ct = ptr if (ct is None or ct[n]) else None
            if l is not None and l[n]:
                for k in range(l[n]):
                    nodes[n].l[k] = ptr
        o = nat.make_opts(kernel=kernel)
        return L.asets_hmf_dev(ndim, 4, 4, nz, nodes, M, its, 0.1, cc, None, C.byref(o))

    cases = [
        (dict(parent=[-1, -1, 0]), "more than one root"),
        (dict(parent=[0, 1]), "bad parent"),
        (dict(parent=[-1, 0, 4]), "bad parent"),
        (dict(parent=[-1, 2, 1]), "no children"),
        (dict(parent=[-1, 0, 3, 2]), "not reachable"),
        (dict(parent=[-1, 0, 0], ct=[0, 1, 0]), "has no Ct"),
        (dict(parent=[-1, 0, 0], l=[0, 1, 0]), "shape directions"),
        (dict(parent=[-1, 0, 0], l=[0, 2, 0], kernel=nat.KERNEL_FUSED), "radius 3"),
        (dict(parent=[-1]), "2 to 4096"),
        (dict(parent=[-1, 0], its=-1), "numIts"),
        (dict(parent=[-1, 0], cc=0.0), "cc"),
        (dict(parent=[-1, 0], nz=2), "nz == 1"),
    ]
    for kw, msg in cases:
        assert call(**kw) == nat.ASETS_EBADARG, kw
        assert msg.encode() in L.asets_last_error(), (kw, L.asets_last_error())
    # 17 levels of internal nodes: a chain root -> ... -> leaf
    chain = [-1] + list(range(17))
    assert call(parent=chain) == nat.ASETS_ETOOBIG
    assert b"deeper" in L.asets_last_error()


def test_cost_builder_validation_before_device_work():
    L = nat.lib()
    buf = np.zeros(64, np.float32)
    ptr = buf.ctypes.data
    ids = (C.c_int32 * 2)(0, 1)
    h = np.ones(9)
    models = (C.c_float * 2)(0.0, 1.0)
    assert L.asets_cost_l1_dev(None, 8, models, 2, ptr, None) == nat.ASETS_EBADARG
    assert L.asets_cost_l1_dev(ptr, 8, models, 65, ptr, None) == nat.ASETS_EBADARG
    assert L.asets_cost_l1_dev(ptr, 0, models, 2, ptr, None) == nat.ASETS_EBADARG
    assert L.asets_cost_smooth_labels_dev(ptr, 4, 4, ids, 2, h.ctypes.data, 4, ptr,
                                          None) == nat.ASETS_EBADARG
    assert b"odd" in L.asets_last_error()
    assert L.asets_cost_smooth_labels_dev(ptr, 4, 4, ids, 2, None, 3, ptr,
                                          None) == nat.ASETS_EBADARG
    assert L.asets_cost_smooth_labels_dev(ptr, 4, 0, ids, 2, h.ctypes.data, 3, ptr,
                                          None) == nat.ASETS_EBADARG
    assert L.asets_cost_smooth_labels_dev(ptr, 4, 4, ids, 0, h.ctypes.data, 3, ptr,
                                          None) == nat.ASETS_EBADARG
