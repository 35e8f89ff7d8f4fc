"""MAT-file (version 5) loading for the reference's data files (data/*.mat).

The reference's applications read their inputs with MATLAB's load():
applications/interactiveGraphCutSegmentation3D.m:31 and smoothManualSegmentations2D.m:31
(brain_1125.mat: img single 91x109x91, scribbles and man_s40_flawed uint16),
atlasProbabilityMapsRegularization3D.m:32 (brain_prob_MNI.mat: bg, csf, wm, gm single).
load_mat reads such files with scipy.io (a parser of the MAT v5 format; nothing in the
file is executed) and returns numeric arrays in MATLAB's shape and column-major layout,
ready for the solvers' column-major convention.
"""
from __future__ import annotations

import numpy as np


def load_mat(path, names=None) -> dict:
    """Numeric variables of a MAT v5 file: {name: ndarray (Fortran order, MATLAB class)}.

    names: variable names to read (None: all).  Raises ValueError for MAT v7.3 (HDF5)
    files, for a requested name that is absent, and for non-numeric variables (cells,
    structs, objects, strings)."""
    import scipy.io

    if isinstance(names, str):
        names = [names]
    try:
        d = scipy.io.loadmat(str(path), variable_names=names, squeeze_me=False,
                             mat_dtype=False, struct_as_record=True)
    except NotImplementedError as e:   # scipy's message for v7.3 / HDF5 files
        raise ValueError(f"{path}: MAT v7.3 (HDF5) is not supported; save with -v7") from e
    out = {}
    for k, v in d.items():
        if k.startswith("__"):
            continue
        if not isinstance(v, np.ndarray) or v.dtype.kind not in "biuf":
            raise ValueError(f"{path}: variable {k!r} is not a numeric array")
        out[k] = np.asfortranarray(v)
    if names is not None:
        missing = [n for n in names if n not in out]
        if missing:
            raise ValueError(f"{path}: no variable(s) {missing}")
    return out


def whos(path) -> list:
    """(name, shape, MATLAB class) of every variable, without reading the data."""
    import scipy.io

    return [(n, tuple(s), c) for n, s, c in scipy.io.whosmat(str(path))]
