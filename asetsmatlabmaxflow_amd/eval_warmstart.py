"""Evaluate the learned warm start against the cold start on 256^3 binary volumes.

    python -m asetsmatlabmaxflow_amd.eval_warmstart --weights w.npz --seeds 5000 5001

For each volume (workloads.binary3d_ellipsoids, the given seeds), with inputs resident on
the GPU:
  reference  plain solver, --ref-iters fixed iterations: E* and the reference labels
  cold       plain solver from the initial cut to errBound (the reference's rule)
  warm       network inference + plain solver from (u0, pt0) to errBound; with
             --variant s2d, inference of the full state + asets_binary_dev_resume
  ideal      plain solver from the reference's own (u, pt) to errBound: the bound on what
             any (u0, pt0) predictor can gain (the flows restart from 0 either way)
  ideal_full_state  (--variant s2d) resume from the reference's whole state
With --sensitivity it instead resumes from the reference's converged state with its face
flows perturbed (noise, scaling, truncation away from the label boundary, blurring,
coarsening) and reports the iterations each start needs.
Reported per run: wall seconds (the median of 7 calls after an untimed one; warm includes
inference), iterations, relative energy gap
(E - E*) / |E*| and label agreement with the reference; for the network output alone its
energy gap and label agreement.  Also the iterations each start needs to reach a relative
energy gap of 1e-5 (searched over a doubling schedule of fixed iteration counts).
One JSON line per volume and a summary line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def _sync():
    import torch

    torch.cuda.synchronize()


def evaluate_volume(net, seed: int, size: int, err_bound: float, ref_iters: int,
                    gap_target: float = 1e-5, gap: bool = True):
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    shape = (size, size, size)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=seed)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    ref = solver.binary(*d, shape, ref_iters, 0.0, cc, st, want_state=True)
    Estar = solver.energy_binary(*d, ref.u, shape)[0]
    lab_ref = ref.u > 0.5

    def quality(u):
        e = solver.energy_binary(*d, u, shape)[0]
        return {"gap": (e - Estar) / abs(Estar),
                "label_agreement": float(((u > 0.5) == lab_ref).float().mean())}

    def timed(fn, n=7):
        # one untimed call (the GPU idled while the volume was generated on the CPU),
        # then the median of n
        fn()
        ts = []
        for _ in range(n):
            _sync()
            t = time.perf_counter()
            r = fn()
            _sync()
            ts.append(time.perf_counter() - t)
        return r, sorted(ts)[n // 2]

    out = {"seed": seed, "shape": list(shape), "E_star": Estar, "ref_iters": ref_iters}
    # warm-up of both paths (kernel loading, workspace allocation)
    solver.binary(*d, shape, 3, 0.0, cc, st)
    if net is not None:
        net.forward(d[0], d[1], d[2], shape, 1)
    r, s = timed(lambda: solver.binary(*d, shape, 20000, err_bound, cc, st))
    out["cold"] = {"seconds": s, "iterations": r.iterations, **quality(r.u)}
    u_id, pt_id = ref.u.clone(), ref.state["pt"].clone()
    r, s = timed(lambda: solver.binary(*d, shape, 20000, err_bound, cc, st, u0=u_id, pt0=pt_id))
    out["ideal_warm"] = {"seconds": s, "iterations": r.iterations, **quality(r.u)}
    full = hasattr(net, "predict_state")
    if full:
        # the exact full state of the reference solve: the bound on what a full-state
        # predictor can gain
        ref_st = {"u": ref.u, **ref.state}
        r, s = timed(lambda: solver.binary_resume(*d, shape, ref_st, 20000, err_bound, cc, st))
        out["ideal_full_state"] = {"seconds": s, "iterations": r.iterations, **quality(r.u)}
    if net is not None:
        if full:
            # inference: the network alone (the resume pairs ps and pt); quality of its
            # output with u clamped to [0, 1], as the resume takes it
            _, s_inf = timed(lambda: net.forward(d[0], d[1], d[2], shape, 1))
            pst = net.predict_state(d[0], d[1], d[2], shape, 1)
            out["network_output"] = {"inference_seconds": s_inf, **quality(pst["u"])}

            def warm():   # u and the faces; the resume pairs ps and pt to them
                w = net.forward(d[0], d[1], d[2], shape, 1)
                return solver.binary_resume(*d, shape, w, 20000, err_bound, cc, st)
        else:
            (u0, pt0), s_inf = timed(lambda: net.forward(d[0], d[1], d[2], shape, 1))
            out["network_output"] = {"inference_seconds": s_inf, **quality(u0)}

            def warm():
                w0, w1 = net.forward(d[0], d[1], d[2], shape, 1)
                return solver.binary(*d, shape, 20000, err_bound, cc, st, u0=w0, pt0=w1)

        r, s = timed(warm)
        out["warm"] = {"seconds": s, "iterations": r.iterations, **quality(r.u)}

    # iterations to a relative energy gap of gap_target (doubling schedule)
    def iters_to_gap(**kw):
        k = 8
        while k <= ref_iters // 2:
            r = solver.binary(*d, shape, k, 0.0, cc, st, **kw)
            if quality(r.u)["gap"] <= gap_target:
                return k
            k *= 2
        return None

    if not gap:
        return out
    out[f"iters_to_gap_{gap_target:g}"] = {"cold": iters_to_gap(),
                                           "ideal_warm": iters_to_gap(u0=u_id, pt0=pt_id)}
    if net is not None and not full:
        u0, pt0 = net.forward(d[0], d[1], d[2], shape, 1)
        out[f"iters_to_gap_{gap_target:g}"]["warm"] = iters_to_gap(u0=u0.clone(),
                                                                   pt0=pt0.clone())
    return out


def sensitivity(seed: int, size: int, err_bound: float, ref_iters: int, calls: int = 3):
    """How accurate a full-state start must be: resume the plain solver from the
    reference's converged state with its face flows perturbed in stated ways (ps and pt
    re-paired to them each time, as for the network's prediction) and report the
    iterations to err_bound and the wall time (median of `calls` calls)."""
    import torch
    import torch.nn.functional as F

    from asetsmatlabmaxflow_amd import solver, workloads
    from asetsmatlabmaxflow_amd.warmstart import pair_state

    shape = (size, size, size)
    Cs, Ct, a, _ = workloads.binary3d_ellipsoids(shape, seed=seed)
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    ref = solver.binary(*d, shape, ref_iters, 0.0, cc, st, want_state=True)
    cold = solver.binary(*d, shape, 20000, err_bound, cc, st)
    gen = torch.Generator(device=d[0].device).manual_seed(seed)
    vol = (size, size, size)   # (nz, nx, ny) view of a column-major field
    faces = {k: ref.state[k] for k in ("bx", "by", "bz")}
    alpha = d[2]

    def noise():
        return torch.randn(ref.u.numel(), generator=gen, device=ref.u.device)

    def blur(f, sigma):
        r = int(3 * sigma + 0.5)
        x = torch.arange(-r, r + 1, device=f.device, dtype=torch.float32)
        k = torch.exp(-0.5 * (x / sigma) ** 2)
        k = k / k.sum()
        t = f.view(1, 1, *vol)
        for dim in range(3):
            shp = [1, 1, 1, 1, 1]
            shp[2 + dim] = 2 * r + 1
            pad = [0, 0, 0, 0, 0, 0]
            pad[2 * (2 - dim)] = pad[2 * (2 - dim) + 1] = r
            t = F.conv3d(F.pad(t, pad, mode="replicate"), k.view(shp))
        return t.reshape(-1)

    def coarse(f):
        t = F.avg_pool3d(f.view(1, 1, *vol), 2)
        return F.interpolate(t, scale_factor=2, mode="nearest").reshape(-1)

    def near_boundary(u, r):
        lab = (u > 0.5).float().view(1, 1, *vol)
        edge = (F.max_pool3d(lab, 3, 1, 1) != -F.max_pool3d(-lab, 3, 1, 1)).float()
        return F.max_pool3d(edge, 2 * r + 1, 1, r).reshape(-1) > 0

    def case(name, fx, u=None):
        state = {"u": (ref.u if u is None else u).clone(),
                 **{k: fx(k, v).contiguous().clone() for k, v in faces.items()}}
        pair_state(state, d[0], d[1], shape)
        r = None
        ts = []
        for _ in range(calls):
            _sync()
            t = time.perf_counter()
            r = solver.binary_resume(*d, shape, state, 20000, err_bound, cc, st)
            _sync()
            ts.append(time.perf_counter() - t)
        return {"start": name, "iterations": r.iterations, "seconds": sorted(ts)[calls // 2]}

    rows = [{"start": "cold (reference initial cut)", "iterations": cold.iterations}]
    rows.append(case("exact", lambda k, v: v))
    for eps in (0.05, 0.1, 0.2, 0.3):
        rows.append(case(f"x (1 + {eps:g} N(0,1))", lambda k, v, e=eps: v * (1 + e * noise())))
    for dl in (0.01, 0.03, 0.1):
        rows.append(case(f"+ {dl:g} alpha N(0,1)", lambda k, v, e=dl: v + e * alpha * noise()))
    for sc in (0.9, 1.1):
        rows.append(case(f"x {sc:g}", lambda k, v, c=sc: v * c))
    for r in (8, 4, 2):
        m = near_boundary(ref.u, r)
        rows.append(case(f"exact within {r} voxels of the label boundary, 0 beyond",
                         lambda k, v, mm=m: torch.where(mm, v, torch.zeros_like(v))))
    for sg in (0.7, 1.5):
        rows.append(case(f"Gaussian blur sigma {sg:g}", lambda k, v, g=sg: blur(v, g)))
    rows.append(case("2x coarser field, upsampled", lambda k, v: coarse(v)))
    rows.append(case("exact faces, u binarised", lambda k, v: v, u=(ref.u > 0.5).float()))
    rows.append(case("exact faces, u blurred sigma 0.7", lambda k, v: v, u=blur(ref.u, 0.7)))
    return {"seed": seed, "shape": list(shape), "err_bound": err_bound, "ref_iters": ref_iters,
            "sensitivity": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=None)
    ap.add_argument("--seeds", type=int, nargs="+", default=[5000, 5001])
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--err-bound", type=float, default=1e-4)
    ap.add_argument("--ref-iters", type=int, default=2000)
    ap.add_argument("--variant", default="full", choices=["full", "s2d"])
    ap.add_argument("--no-gap", action="store_true",
                    help="skip the iterations-to-energy-gap search")
    ap.add_argument("--sensitivity", action="store_true",
                    help="instead: resume from the reference's state with perturbed flows")
    args = ap.parse_args()
    if args.sensitivity:
        for seed in args.seeds:
            print(json.dumps(sensitivity(seed, args.size, args.err_bound, args.ref_iters)),
                  flush=True)
        return

    import torch

    from asetsmatlabmaxflow_amd.warmstart import WarmStartNet, WarmStartNetS2D

    net = None
    if args.weights:
        net = (WarmStartNetS2D if args.variant == "s2d" else WarmStartNet)(
            torch.device("cuda:0"))
        net.load(args.weights)
    rows = []
    for seed in args.seeds:
        row = evaluate_volume(net, seed, args.size, args.err_bound, args.ref_iters,
                              gap=not args.no_gap)
        rows.append(row)
        print(json.dumps(row), flush=True)
    keys = ["cold", "ideal_warm"] + (["ideal_full_state"] if "ideal_full_state" in rows[0]
                                     else []) + (["warm"] if net is not None else [])
    summary = {k: {"mean_seconds": float(np.mean([r[k]["seconds"] for r in rows])),
                   "mean_iterations": float(np.mean([r[k]["iterations"] for r in rows])),
                   "mean_gap": float(np.mean([r[k]["gap"] for r in rows])),
                   "mean_label_agreement": float(np.mean([r[k]["label_agreement"]
                                                          for r in rows]))}
               for k in keys}
    if net is not None:
        summary["speedup_warm_vs_cold"] = summary["cold"]["mean_seconds"] / \
            summary["warm"]["mean_seconds"]
    print(json.dumps({"summary": summary, "volumes": len(rows), "size": args.size,
                      "err_bound": args.err_bound}), flush=True)


if __name__ == "__main__":
    main()
