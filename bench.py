"""Benchmark: 256^3 binary continuous max-flow on one MI355X.

Workload: the 256x256x256 binary volume of asetsmatlabmaxflow_amd.workloads
.binary3d_ellipsoids (seed 1002; 12 ellipsoids, Gaussian NLL costs, contrast-sensitive
alpha), cc = 0.35, steps = 0.11.  A step is one solve of 200 fixed iterations
(errBound 0) on inputs already resident in HBM; value = voxel-iterations per second over
the K timed steps.  After the timed steps the same volume is solved to errBound 1e-4
(time-to-tolerance).  The CPU baseline is the single-threaded FP32 reference-order oracle
timed on a bounded number of iterations of the same volume.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_VOXEL_ITER = 60  # Cs, Ct, alpha read; u, ps, pt, bx, by, bz read + written (fp32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--kernel", default="fused", choices=["fused", "passes"])
    ap.add_argument("--tile-x", type=int, default=0)
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--cpu-iters", type=int, default=int(os.environ.get("ASETS_CPU_ITERS", 20)))
    ap.add_argument("--no-tol", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    n = args.size
    shape = (n, n, n)
    N = n ** 3
    t0 = time.time()
    Cs, Ct, alpha, _ = workloads.binary3d_ellipsoids(shape, seed=1002)
    gen_s = time.time() - t0
    dev = torch.device("cuda:0")
    dCs, dCt, dA = (solver.to_device(x, dev) for x in (Cs, Ct, alpha))
    cc, steps = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    stream = torch.cuda.Stream()
    u_out = torch.empty(N, dtype=torch.float32, device=dev)

    def solve(max_iter, eb, time_kernel=False):
        with torch.cuda.stream(stream):
            return solver.binary(dCs, dCt, dA, shape, max_iter, eb, cc, steps,
                                 kernel=args.kernel, tile_x=args.tile_x, zchunk=args.zchunk,
                                 stream=stream, time_kernel=time_kernel, out=u_out)

    for _ in range(args.warmup):
        solve(args.iters, 0.0)
    torch.cuda.synchronize()
    kms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = solve(args.iters, 0.0, time_kernel=True)
        kms.append(r.kernel_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall * 1e3 / args.steps
    value = N * args.iters * args.steps / wall
    k_ms = float(np.mean(kms))
    achieved = BYTES_PER_VOXEL_ITER * N / (k_ms * 1e-3) / 1e9
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs of
    # this kernel and workload (scaled by a known-size copy), committed under profiles/
    traffic = None
    prof = ROOT / "profiles" / "kf_binary_256cubed.json"
    if prof.exists() and n == 256 and args.kernel == "fused" and not args.tile_x \
            and not args.zchunk:
        traffic = json.loads(prof.read_text())["traffic_bytes_per_launch"]

    out = {
        "metric": "voxel-iters/s, 256^3 binary MF (plain, 200 fixed iterations); "
                  "time-to-tol at errBound 1e-4; HBM GB/s",
        "value": value,
        "unit": "voxel-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid volume, workloads.binary3d_ellipsoids seed 1002)",
        "config": {"workload": f"binary3d {n}^3 ellipsoids, contrast alpha, cc 0.35 steps 0.11",
                   "shape": list(shape), "iters_per_step": args.iters, "kernel": args.kernel},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel_ms_per_iter": k_ms,
                     "bytes_per_voxel_iter": BYTES_PER_VOXEL_ITER},
        "gen_s": gen_s,
    }

    # fp16 flow storage (ps, pt, bx, by, bz in fp16; u and arithmetic fp32): 40 B/voxel-iter
    if not args.no_fp16:
        with torch.cuda.stream(stream):
            solver.binary(dCs, dCt, dA, shape, 20, 0.0, cc, steps, flow="f16", stream=stream)
            r16 = solver.binary(dCs, dCt, dA, shape, args.iters, 0.0, cc, steps, flow="f16",
                                stream=stream, time_kernel=True)
        torch.cuda.synchronize()
        out["fp16_flows"] = {"kernel_ms_per_iter": r16.kernel_ms,
                             "voxel_iter_per_s": N / (r16.kernel_ms * 1e-3),
                             "bytes_per_voxel_iter": 40,
                             "achieved_GBps": 40 * N / (r16.kernel_ms * 1e-3) / 1e9,
                             "max_abs_u_diff_vs_fp32": float((r16.u - u_out).abs().max().item())}

    if not args.no_tol:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rt = solve(20000, 1e-4)
        torch.cuda.synchronize()
        ttt = time.perf_counter() - t1
        lab = (rt.u > 0.5)
        out["time_to_tol"] = {"err_bound": 1e-4, "seconds": ttt, "it_num": rt.it_num,
                              "iterations": rt.iterations,
                              "final_cvg": float(rt.cvg[rt.iterations - 1].item()),
                              "fg_fraction": float(lab.float().mean().item())}

    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as orc

        k = max(1, args.cpu_iters)
        t2 = time.perf_counter()
        orc.binary(Cs, Ct, alpha, shape, k, 0.0, cc, steps)
        cpu_s = time.perf_counter() - t2
        out["cpu_baseline"] = {"value": N * k / cpu_s, "unit": "voxel-iter/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations of the same {n}^3 volume "
                                         f"(FP32 reference-order oracle, gcc -O2, "
                                         f"1 thread), {cpu_s:.1f} s incl. init"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
