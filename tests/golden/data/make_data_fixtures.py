"""Writes tests/golden/data/brain_crops.npz from the reference's data files (data/*.mat, MAT
v5): crops of the real inputs of the application scripts, small enough to commit, for
the GPU tests of asetsmatlabmaxflow_amd.applications (which run where the reference is
not present).  Inputs only -- the reference stores no outputs for them.

  img3, scribbles3          brain_1125.mat img / scribbles, all 91 rows, cols 16:96,
                            slices 41:49 (0-based; every scribble lies on slice 45, the
                            background ones along the image border)
  bg3, csf3, wm3, gm3       brain_prob_MNI.mat, the same crop
  img2, man2                brain_1125.mat img and man_s40_flawed, slice 40 (1-based)

Run from the repository root: python tests/golden/data/make_data_fixtures.py
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[3]
sys.path.insert(0, str(ROOT))
from asetsmatlabmaxflow_amd import matdata  # noqa: E402

DATA = Path("/root/reference/data")
CROP = (slice(0, 91), slice(16, 96), slice(41, 49))


def main():
    b = matdata.load_mat(DATA / "brain_1125.mat", ["img", "scribbles", "man_s40_flawed"])
    p = matdata.load_mat(DATA / "brain_prob_MNI.mat", ["bg", "csf", "wm", "gm"])
    out = {
        "img3": b["img"][CROP], "scribbles3": b["scribbles"][CROP],
        "bg3": p["bg"][CROP], "csf3": p["csf"][CROP], "wm3": p["wm"][CROP],
        "gm3": p["gm"][CROP],
        "img2": b["img"][:, :, 39], "man2": b["man_s40_flawed"][:, :, 39],
    }
    for l in range(1, 5):
        assert (out["scribbles3"] == l).any(), f"label {l} not scribbled in the crop"
    out = {k: np.ascontiguousarray(v) for k, v in out.items()}
    dst = Path(__file__).resolve().parent / "brain_crops.npz"
    np.savez_compressed(dst, **out)
    print(dst, dst.stat().st_size, {k: (v.dtype.name, v.shape) for k, v in out.items()})


if __name__ == "__main__":
    main()
