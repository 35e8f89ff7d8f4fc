"""Build the native parts in-tree (no JIT cache: the built files travel with the tree).

Products (next to this file):
  libasets_hip.so            C-ABI HIP library (include/asets.h), gfx950 code objects
  libasets_nn.so             warm-start network kernels (include/asets_nn.h)
  mex/libmexshim.so          project-written minimal mx/mex API (tests, Python driver)
  mex/asets*_mex.so          the six mex-ABI host files, linked against the two above

`python -m asetsmatlabmaxflow_amd.build` builds everything; `build_oracle()` builds the
CPU checker under oracle/ (test infrastructure, not part of the product).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
MEX_OUT = PKG / "mex"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ASETS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no a*b+c contraction (bit parity with the reference order).
# -packed-fp32-ops disabled: no v_pk_{add,mul,fma}_f32 generation.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c99", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-Wno-unused-function"]

MEX_SOURCES = ["asetsBinaryMF2D_mex", "asetsBinaryMF3D_mex", "asetsPotts2D_mex",
               "asetsPotts3D_mex", "asetsIshikawa2D_mex", "asetsIshikawa3D_mex"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{msg}")


def _atomic_out(path: Path) -> Path:
    return path.with_name(path.name + ".tmp")


def build_hip_library() -> Path:
    out = PKG / "libasets_hip.so"
    tmp = _atomic_out(out)
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}",
          "-o", str(tmp), str(CSRC / "asets_hip.hip")])
    os.replace(tmp, out)
    return out


def build_nn_library() -> Path:
    out = PKG / "libasets_nn.so"
    tmp = _atomic_out(out)
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INCLUDE}", f"-I{CSRC}",
          "-o", str(tmp), str(CSRC / "nn_kernels.hip")])
    os.replace(tmp, out)
    return out


def build_mex() -> list[Path]:
    MEX_OUT.mkdir(exist_ok=True)
    shim_dir = CSRC / "mex" / "shim"
    shim = MEX_OUT / "libmexshim.so"
    tmp = _atomic_out(shim)
    _run(["gcc", *C_FLAGS, "-shared", f"-I{shim_dir}", "-o", str(tmp),
          str(shim_dir / "mexshim.c")])
    os.replace(tmp, shim)
    outs = [shim]
    for name in MEX_SOURCES:
        so = MEX_OUT / f"{name}.so"
        tmp = _atomic_out(so)
        _run(["gcc", *C_FLAGS, "-shared", f"-I{shim_dir}", f"-I{INCLUDE}",
              f"-I{CSRC / 'mex'}", "-o", str(tmp), str(CSRC / "mex" / f"{name}.c"),
              f"-L{MEX_OUT}", "-lmexshim", f"-L{PKG}", "-lasets_hip",
              "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."])
        os.replace(tmp, so)
        outs.append(so)
    return outs


def build_oracle() -> Path:
    odir = ROOT / "oracle"
    _run(["make", "-s", "-C", str(odir), "CC=gcc"])
    return odir / "liboracle.so"


def build_all(oracle: bool = True) -> list[Path]:
    outs = [build_hip_library(), build_nn_library()]
    outs += build_mex()
    if oracle:
        outs.append(build_oracle())
    return outs


if __name__ == "__main__":
    for p in build_all(oracle="--no-oracle" not in sys.argv):
        print(p)
