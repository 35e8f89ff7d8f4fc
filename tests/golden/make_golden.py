"""Writes tests/golden/*.npz: small input/output vectors of the FP64 linear-algebra
restatement (tests/matrix_form.py), against which the oracle is pinned.

The reference's own C sources cannot be built in this image (they include MATLAB's
mex.h) and its tests keep no stored outputs, so these fixtures come from the
independent restatement, with seeded inputs in the manner of the reference tests
(tests/runTestBinaryMF.m:36-44, tests/runTestPotts.m:37-46: Gaussian-smoothed uniform
costs, constant or per-label alpha).

Run:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

import matrix_form as mf  # noqa: E402
from asetsmatlabmaxflow_amd import workloads  # noqa: E402

CASES_BINARY = [
    # name, shape, alpha, iters, cc, steps
    ("bin2d_17x13", (17, 13), 0.25, 25, 0.2, 0.16),
    ("bin2d_9x31", (9, 31), 0.05, 40, 0.3, 0.16),
    ("bin3d_5x7x3", (5, 7, 3), 0.1, 20, 0.25, 0.11),
    ("bin3d_11x7x5", (11, 7, 5), 0.02, 30, 0.35, 0.11),
]
CASES_POTTS = [
    ("potts2d_13x11_L3", (13, 11), 3, 0.05, 20, 0.2, 0.16),
    ("potts3d_9x7x5_L3", (9, 7, 5), 3, 0.025, 15, 0.25, 0.11),
    ("potts3d_6x5x4_L5", (6, 5, 4), 5, 0.02, 12, 0.25, 0.11),
]


def main():
    for i, (name, shape, a, iters, cc, steps) in enumerate(CASES_BINARY):
        Cs, Ct, alpha = workloads.random_smooth_binary(shape, seed=100 + i, alpha=a)
        u, ps, pt, b, cvg = mf.binary(Cs, Ct, alpha, shape, iters, cc, steps)
        np.savez_compressed(HERE / f"{name}.npz", Cs=Cs.ravel(order="F"),
                            Ct=Ct.ravel(order="F"), alpha=alpha.ravel(order="F"),
                            shape=np.array(shape), iters=iters, cc=cc, steps=steps, u=u, ps=ps,
                            pt=pt, cvg=cvg)
    for i, (name, shape, L, astep, iters, cc, steps) in enumerate(CASES_POTTS):
        Ct, alpha = workloads.random_smooth_potts(shape, L, seed=200 + i, alpha_step=astep)
        u, ps, cvg = mf.potts(Ct, alpha, shape, L, iters, cc, steps)
        np.savez_compressed(HERE / f"{name}.npz", Ct=Ct.ravel(order="F"),
                            alpha=alpha.ravel(order="F"), shape=np.array(shape), nlab=L,
                            iters=iters, cc=cc, steps=steps, u=u, ps=ps, cvg=cvg)
    print("wrote", len(CASES_BINARY) + len(CASES_POTTS), "fixtures")


if __name__ == "__main__":
    main()
