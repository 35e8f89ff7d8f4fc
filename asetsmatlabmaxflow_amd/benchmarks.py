"""Secondary benchmarks (one JSON line each); the headline benchmark is bench.py.

    python -m asetsmatlabmaxflow_amd.benchmarks config1   # 2D 256^2 ellipses, 300 iterations
    python -m asetsmatlabmaxflow_amd.benchmarks config3   # Potts 192^3, L = 6, 300 iterations
    python -m asetsmatlabmaxflow_amd.benchmarks config4   # 64 uneven binary volumes, batched,
                                                         # to errBound 1e-4, fp32 vs fp16 flows

Algorithmic bytes per voxel-iteration: binary 2D 52, binary 3D 60 (fp32) / 40 (fp16
flows), Potts 3D 48L+8.  The CPU figures come from the FP32 reference-order oracle under
oracle/ (single thread), timed on a bounded number of iterations.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
HBM_PEAK_GBS = 8000.0


def _oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    return orc


def config1(args):
    import torch

    from . import solver, workloads

    Cs, Ct, a, pars, _ = workloads.binary2d_ellipses(256, seed=1001)
    shape = (256, 256)
    N = 256 * 256
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    solver.binary(*d, shape, 300, 0.0, 0.3, 0.16)
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = solver.binary(*d, shape, 300, 0.0, 0.3, 0.16, time_kernel=True)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t
    orc = _oracle()
    t = time.perf_counter()
    ref = orc.binary(Cs, Ct, a, shape, 300, 0.0, 0.3, 0.16)
    cpu_s = time.perf_counter() - t
    same = bool(np.array_equal(r.u.cpu().numpy().view(np.uint32), ref.u.view(np.uint32)))
    return {"benchmark": "config1 binary2d 256^2, 300 fixed iterations",
            "gpu_seconds": gpu_s, "gpu_ms_per_iter": r.kernel_ms,
            "gpu_voxel_iter_per_s": N * 300 / gpu_s,
            "cpu_seconds": cpu_s, "cpu_voxel_iter_per_s": N * 300 / cpu_s, "cpu_cores": 1,
            "u_bit_identical_to_oracle": same}


def config3(args):
    import torch

    from . import solver, workloads

    shape, L = (192, 192, 192), 6
    Ct, a, _ = workloads.potts3d_voronoi(shape, nlab=L, seed=1003)
    N = int(np.prod(shape))
    d = [solver.to_device(x) for x in (Ct, a)]
    cc, st = workloads.POTTS3D_CC, workloads.POTTS3D_STEPS
    out = {"benchmark": "config3 potts3d 192^3 L=6, 300 fixed iterations"}
    for kern in ("fused", "passes"):
        solver.potts(*d, shape, L, 5, 0.0, cc, st, kernel=kern)
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = solver.potts(*d, shape, L, 300, 0.0, cc, st, kernel=kern, time_kernel=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        gbs = (48 * L + 8) * N / (r.kernel_ms * 1e-3) / 1e9
        out[kern] = {"seconds": wall, "ms_per_iter": r.kernel_ms,
                     "voxel_iter_per_s": N * 300 / wall, "achieved_GBps": gbs,
                     "roofline_frac": gbs / HBM_PEAK_GBS}
        if kern == "fused":
            u_f = r.u.clone()
            lab = u_f.view(L, -1).argmax(0)
    k = max(1, args.cpu_iters)
    orc = _oracle()
    t = time.perf_counter()
    orc.potts(Ct, a, shape, L, k, 0.0, cc, st)
    cpu_s = time.perf_counter() - t
    out["cpu_baseline"] = {"voxel_iter_per_s": N * k / cpu_s, "cores": 1,
                           "sample": f"{k} iterations incl. init, FP32 reference-order oracle"}
    out["label_fractions"] = [float((lab == l).float().mean()) for l in range(L)]
    return out


def config4(args):
    import torch

    from . import solver, workloads

    rng = np.random.default_rng(1004)
    sizes = [tuple(int(v) for v in rng.integers(64, 193, 3)) for _ in range(args.nvol)]
    host, dev = [], []
    t = time.perf_counter()
    for v, sh in enumerate(sizes):
        Cs, Ct, a, _ = workloads.binary3d_ellipsoids(sh, seed=1004 * 1000 + v)
        host.append((Cs, Ct, a, sh))
        dev.append(tuple(solver.to_device(x) for x in (Cs, Ct, a)) + (sh,))
    gen_s = time.perf_counter() - t
    total = int(sum(np.prod(s) for s in sizes))
    out = {"benchmark": f"config4 batched: {args.nvol} binary volumes, axes U{{64..192}}, "
                        f"errBound 1e-4", "volumes": args.nvol, "total_voxels": total,
           "gen_s": gen_s}
    cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    res = {}
    for flow in ("f32", "f16"):
        solver.binary_batched(dev, 3, 0.0, cc, st, flow=flow)   # tuning + warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        rs, secs, loop_ms = solver.binary_batched(dev, 2000, 1e-4, cc, st, flow=flow,
                                                  time_kernel=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        vit = sum(int(np.prod(s)) * r.iterations for s, r in zip(sizes, rs))
        bpv = 60 if flow == "f32" else 40
        res[flow] = rs
        out[flow] = {"seconds": wall, "loop_ms": loop_ms, "voxel_iterations": vit,
                     "voxel_iter_per_s": vit / wall,
                     "achieved_GBps_loop": bpv * vit / (loop_ms * 1e-3) / 1e9,
                     "iterations": [r.iterations for r in rs]}
    # error against the FP64 oracle on a few volumes, same iteration counts as fp32
    orc = _oracle()
    errs = []
    for v in range(min(args.check, args.nvol)):
        Cs, Ct, a, sh = host[v]
        it = res["f32"][v].iterations
        ref = orc.binary(Cs, Ct, a, sh, it, 0.0, cc, st, precision="f64")
        e = {}
        for flow in ("f32", "f16"):
            u = res[flow][v].u.cpu().numpy().astype(np.float64)
            e[flow] = {"max_abs_u_err": float(np.abs(u - ref.u).max()),
                       "label_agreement": float(((u > 0.5) == (ref.u > 0.5)).mean()),
                       "iterations": res[flow][v].iterations}
        errs.append({"volume": v, "shape": sh, **e})
    out["error_vs_fp64_oracle"] = errs
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["config1", "config3", "config4"])
    ap.add_argument("--cpu-iters", type=int, default=3)
    ap.add_argument("--nvol", type=int, default=64)
    ap.add_argument("--check", type=int, default=3)
    args = ap.parse_args()
    r = {"config1": config1, "config3": config3, "config4": config4}[args.which](args)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
