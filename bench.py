"""Benchmark of the continuous max-flow solvers on one MI355X.  Prints one JSON line.

Default workload (--workload binary3d): the 256x256x256 binary volume of
asetsmatlabmaxflow_amd.workloads.binary3d_ellipsoids (seed 1002; 12 ellipsoids, Gaussian
NLL costs, contrast-sensitive alpha), cc = 0.35, steps = 0.11.  A step is one solve of 200
fixed iterations (errBound 0) on inputs already resident in HBM; value = voxel-iterations
per second over the K timed steps.  After the timed steps the same volume is solved to
errBound 1e-4 (time-to-tolerance), with fp16 flow storage, and warm-started by the
learned network when its weights are present (asetsmatlabmaxflow_amd/weights).  The CPU
baseline is the single-threaded FP32 reference-order oracle timed on a bounded number of
iterations of the same volume.

Other workloads (one JSON line each, same keys):
  binary2d  256x256 ellipses (seed 1001), 300 fixed iterations, cc 0.3, steps 0.16
  potts3d   192^3, L = 6 Voronoi labels (seed 1003), 300 fixed iterations, cc 0.25
  batched   64 binary volumes with axes uniform in {64..192} (seeds 1004*1000+v), one
            graph-replayed persistent launch per iteration, to errBound 1e-4, fp32 and
            fp16 flow storage
  ishikawa3d  192^3, L = 6 ordered labels (seed 1003), 300 fixed iterations, cc 0.75
  hmf2d     256x256 hierarchical max-flow, flat tree of 40 leaves (the linear model of
            tests/runTestHMFLinearModel.m, seed 1005), 200 fixed iterations
  hmf3d     192^3 hierarchical max-flow, root -> 2 branches -> 3 leaves (seed 1006), 100
            fixed iterations
  hbm       the HBM ceiling: plain read, copy and solver-traffic-mix streams with
            16-byte accesses over about 1 GB (no solver; the yardstick of the rooflines)

Algorithmic bytes per voxel-iteration: binary 3D 60 (fp32 flows) / 40 (fp16 flows),
binary 2D 52, Potts 3D 48L+8, Ishikawa 3D 12L+36(L-1) (read Ct, alpha; read and write u,
pt and the three face fields), HMF 8 + 60 (3D) or 52 (2D) per non-root node.
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
REF_ITERS = 2000        # fixed iterations of the converged reference solve
LABEL_TARGET = 0.9999   # label agreement the errBound 1e-4 solve is held to
TIMED_CALLS = 9         # whole solve calls timed per figure (after an untimed first call)


def _median_call(fn, n=TIMED_CALLS):
    """Median wall time of n calls of fn() (each bracketed by device synchronisations)
    and the last call's result."""
    import torch

    ts, res = [], None
    for _ in range(n):
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    return sorted(ts)[n // 2], res, ts


def _oracle():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc

    return orc


def _progress(msg):
    """One line on stderr as each phase of a run ends (the JSON line stays alone on
    stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _host_cpu():
    """(model name, logical CPUs visible to this process)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count()
    return model, n


def _cpu_time(fn, repeats):
    """Wall seconds of fn() on one core (this process pinned to its first allowed CPU
    for the duration), median of `repeats` runs."""
    prev = None
    try:
        prev = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(prev)})
    except (AttributeError, OSError):
        prev = None
    try:
        ts = []
        for _ in range(repeats):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
    finally:
        if prev is not None:
            os.sched_setaffinity(0, prev)
    return sorted(ts)[len(ts) // 2], ts


def _base(args, metric, value, unit, ms_per_step, dtype, data, config):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data,
            "config": config}


def _roofline(bytes_per_unit, units, k_ms, traffic=None):
    achieved = bytes_per_unit * units / (k_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel_ms_per_iter": k_ms,
            "bytes_per_voxel_iter": bytes_per_unit}


def bench_binary3d(args):
    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    n = args.size
    shape = (n, n, n)
    N = n ** 3
    t0 = time.time()
    Cs, Ct, alpha, _ = workloads.binary3d_ellipsoids(shape, seed=1002)
    gen_s = time.time() - t0
    _progress(f"volume generated in {gen_s:.1f} s")
    dev = torch.device("cuda:0")
    dCs, dCt, dA = (solver.to_device(x, dev) for x in (Cs, Ct, alpha))
    cc, steps = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    stream = torch.cuda.Stream()
    u_out = torch.empty(N, dtype=torch.float32, device=dev)

    def solve(max_iter, eb, time_kernel=False, **kw):
        with torch.cuda.stream(stream):
            return solver.binary(dCs, dCt, dA, shape, max_iter, eb, cc, steps,
                                 kernel=args.kernel, tile_x=args.tile_x, zchunk=args.zchunk,
                                 stream=stream, time_kernel=time_kernel, **kw)

    for _ in range(args.warmup):
        solve(args.iters, 0.0, out=u_out)
    torch.cuda.synchronize()
    kms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r = solve(args.iters, 0.0, time_kernel=True, out=u_out)
        kms.append(r.kernel_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    k_ms = float(np.mean(kms))
    _progress(f"{args.steps} timed steps: {wall:.3f} s, {k_ms:.4f} ms per iteration")
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs of
    # this kernel and workload (scaled by a known-size copy), committed under profiles/
    traffic = None
    prof = ROOT / "profiles" / "kf_binary_256cubed.json"
    if prof.exists() and n == 256 and args.kernel == "fused" and not args.tile_x \
            and not args.zchunk:
        traffic = json.loads(prof.read_text())["traffic_bytes_per_launch"]
    out = _base(args, "voxel-iters/s, 256^3 binary MF (plain, 200 fixed iterations); "
                      "time-to-tol at errBound 1e-4; HBM GB/s",
                N * args.iters * args.steps / wall, "voxel-iter/s", wall * 1e3 / args.steps,
                "f32", "synthetic (seeded ellipsoid volume, workloads.binary3d_ellipsoids "
                       "seed 1002)",
                {"workload": f"binary3d {n}^3 ellipsoids, contrast alpha, cc 0.35 steps 0.11",
                 "shape": list(shape), "iters_per_step": args.iters, "kernel": args.kernel})
    out["roofline"] = _roofline(60, N, k_ms, traffic)
    out["roofline"]["traffic_source"] = (
        "not measured in this run: the constant of profiles/kf_binary_256cubed.json, from "
        "rocprofv3 FETCH_SIZE and WRITE_SIZE counter runs of this kernel and workload "
        "scaled by a 256 MiB device copy" if traffic is not None else None)
    out["gen_s"] = gen_s

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
        if not args.no_cpu:
            # both flow storages against the FP64 oracle run for the same iterations on
            # the central (n/2)^3 crop of the volume
            orc = _oracle()
            crop = tuple(slice(n // 4, n // 4 + n // 2) for _ in range(3))
            cs = [np.asfortranarray(x[crop]) for x in (Cs, Ct, alpha)]
            csh = cs[0].shape
            dcs = [solver.to_device(x, dev) for x in cs]
            r64 = orc.binary(*cs, csh, args.iters, 0.0, cc, steps, precision="f64")
            err = {}
            for flow in ("f32", "f16"):
                with torch.cuda.stream(stream):
                    rc_ = solver.binary(*dcs, csh, args.iters, 0.0, cc, steps, flow=flow,
                                        stream=stream)
                torch.cuda.synchronize()
                ug = rc_.u.cpu().numpy().astype(np.float64)
                d = np.abs(ug - r64.u)
                err[flow] = {"max_abs_u_err": float(d.max()), "mean_abs_u_err": float(d.mean()),
                             "label_agreement": float(np.mean((ug > 0.5) == (r64.u > 0.5)))}
            out["fp16_flows"]["error_vs_fp64_oracle"] = {
                "crop": "x".join(map(str, csh)), "iterations": args.iters,
                "fp16_flows": err["f16"], "fp32_flows": err["f32"]}
            _progress("fp16 error against the FP64 oracle done")

    if not args.no_tol:
        # whole solve calls (setup, iterations, stop polls, output copy) from the host:
        # the first call of this configuration, then the median of the next TIMED_CALLS.
        # The GPU idled through the CPU oracle runs above: 100 fixed iterations first
        # bring its clocks back up (without them the first calls ran up to 20 % slower)
        solve(100, 0.0)
        first, _, _ = _median_call(lambda: solve(20000, 1e-4), 1)
        ttt, rt, runs = _median_call(lambda: solve(20000, 1e-4))
        runs = [first] + runs
        _progress(f"time to tolerance: {ttt * 1e3:.2f} ms")
        lab = (rt.u > 0.5)
        e = solver.energy_binary(dCs, dCt, dA, rt.u, shape)
        out["time_to_tol"] = {"err_bound": 1e-4, "seconds": ttt, "first_call_seconds": runs[0],
                              "runs_seconds": runs, "it_num": rt.it_num,
                              "iterations": rt.iterations,
                              "final_cvg": float(rt.cvg[rt.iterations - 1].item()),
                              "fg_fraction": float(lab.float().mean().item()),
                              "energy": float(e[0])}
        # labels (u > 0.5) against a converged solve (REF_ITERS fixed iterations), the
        # >= 99.99 % agreement target, and the largest errBound that meets it
        ref_u = solve(REF_ITERS, 0.0).u.clone()
        lab_ref = ref_u > 0.5

        def agree(u):
            return float(((u > 0.5) == lab_ref).float().mean().item())

        sweep = []
        for eb in (1e-4, 5e-5, 2e-5, 1e-5, 5e-6, 2e-6, 1e-6, 5e-7):
            t_eb, r_eb, _ = _median_call(lambda: solve(20000, eb))
            sweep.append({"err_bound": eb, "seconds": t_eb,
                          "iterations": r_eb.iterations, "label_agreement": agree(r_eb.u)})
            if sweep[-1]["label_agreement"] >= LABEL_TARGET:
                break
        first = next((x for x in sweep if x["label_agreement"] >= LABEL_TARGET), None)
        out["time_to_tol"]["labels_vs_converged"] = {
            "reference": f"{REF_ITERS} fixed iterations",
            "label_agreement": agree(rt.u), "target": LABEL_TARGET,
            "meets_target": agree(rt.u) >= LABEL_TARGET,
            "first_err_bound_meeting_target": first["err_bound"] if first else None,
            "time_to_tol_at_that_err_bound_s": first["seconds"] if first else None,
            "iterations_at_that_err_bound": first["iterations"] if first else None,
            "sweep": sweep}
        _progress("label agreement sweep done")
        # certified bracket D(p~) <= min <= P(u) (FP64, asets_binary_bounds_dev) of the
        # errBound-1e-4 state and of the converged one; it levels off instead of closing
        # (the reference's rescale is not a projection), so it is reported, not a stop rule
        bnd = {}
        for label, its, eb in (("err_bound_1e-4", 20000, 1e-4), ("fixed_2000", REF_ITERS, 0.0)):
            rs = solve(its, eb, want_state=True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            b = solver.binary_bounds(dCs, dCt, dA, rs.u, rs.state["bx"], rs.state["by"],
                                     rs.state["bz"], shape, sweeps=30)
            torch.cuda.synchronize()
            bnd[label] = {**b, "iterations": rs.iterations, "seconds": time.perf_counter() - t1}
            del rs
        bnd["note"] = ("primal: P(u) with the regulariser bounded after 30 red-black lambda "
                       "sweeps; primal_tv: the voxel-centred bound; width = (primal - dual) / "
                       "|dual|")
        out["time_to_tol"]["bounds"] = bnd
        _progress("primal-dual bounds done")

    # learned warm starts vs the cold start, both to errBound 1e-4, energies against a
    # 2000-iteration reference solve: "warm_start" = the space-to-depth network's full state
    # (u, faces; ps, pt paired to them) resumed by asets_binary_dev_resume;
    # "warm_start_u_pt" = the full-resolution network's (u0, pt0) with the flows at 0
    wdir = ROOT / "asetsmatlabmaxflow_amd" / "weights"
    variants = (("warm_start", "warmstart_s2d.npz", 16), ("warm_start_u_pt",
                                                          "warmstart_unet.npz", 8))
    if not args.no_tol and not args.no_warm and any((wdir / w).exists() for _, w, _ in variants):
        from asetsmatlabmaxflow_amd.warmstart import WarmStartNet, WarmStartNetS2D

        with torch.cuda.stream(stream):
            ref = solver.binary(dCs, dCt, dA, shape, REF_ITERS, 0.0, cc, steps, stream=stream)
            e_star = solver.energy_binary(dCs, dCt, dA, ref.u, shape, stream=stream)[0]
            lab_ref = ref.u > 0.5
            e_cold = out["time_to_tol"]["energy"]
            lab_cold = float(((rt.u > 0.5) == lab_ref).float().mean().item())
            for key, wname, mult in variants:
                if not (wdir / wname).exists() or n % mult:
                    continue
                full = key == "warm_start"
                net = (WarmStartNetS2D if full else WarmStartNet)(dev)
                net.load(wdir / wname)

                def predict():
                    # full state: u and the faces; the resume pairs ps and pt to them
                    if full:
                        return net.forward(dCs, dCt, dA, shape, 1)
                    u0, pt0 = net.forward(dCs, dCt, dA, shape, 1)
                    return {"u": u0, "pt": pt0}

                def warm_solve(st):
                    if full:
                        return solver.binary_resume(dCs, dCt, dA, shape, st, 20000, 1e-4, cc,
                                                    steps, stream=stream)
                    return solver.binary(dCs, dCt, dA, shape, 20000, 1e-4, cc, steps,
                                         u0=st["u"], pt0=st["pt"], stream=stream)

                warm_solve(predict())   # warm-up
                # medians of TIMED_CALLS calls, as the cold figure: inference alone, and
                # inference + solve
                t_inf, _, _ = _median_call(predict)
                t_warm, rw, _ = _median_call(lambda: warm_solve(predict()))
                e_warm = solver.energy_binary(dCs, dCt, dA, rw.u, shape, stream=stream)[0]
                # the network's own output (u clamped to [0, 1] as the resume takes it)
                st = net.predict_state(dCs, dCt, dA, shape, 1) if full else predict()
                e_net = solver.energy_binary(dCs, dCt, dA, st["u"], shape, stream=stream)[0]
                out[key] = {
                    "start": ("network full state (u, faces; ps, pt paired), resumed"
                              if full else "network (u0, pt0), flows 0"),
                    "err_bound": 1e-4, "inference_seconds": t_inf, "seconds": t_warm,
                    "iterations": rw.iterations, "cold_seconds": out["time_to_tol"]["seconds"],
                    "cold_iterations": out["time_to_tol"]["iterations"],
                    "speedup_vs_cold": out["time_to_tol"]["seconds"] / t_warm,
                    "energy_gap_warm": (e_warm - e_star) / abs(e_star),
                    "energy_gap_cold": (e_cold - e_star) / abs(e_star),
                    "energy_gap_network_output": (e_net - e_star) / abs(e_star),
                    "label_agreement_warm": float(((rw.u > 0.5) == lab_ref).float().mean()
                                                  .item()),
                    "label_agreement_cold": lab_cold,
                    "label_agreement_network_output": float(((st["u"] > 0.5) == lab_ref)
                                                            .float().mean().item()),
                    "weights": f"asetsmatlabmaxflow_amd/weights/{wname}"}
                # the same final-energy accuracy: the loosest errBound of a fixed schedule at
                # which the warm start's energy gap is no worse than the cold 1e-4 stop's
                gap_cold = (e_cold - e_star) / abs(e_star)
                for eb in (1e-4, 5e-5, 3e-5, 2e-5, 1.5e-5, 1e-5, 5e-6):
                    def warm_eb():
                        st = predict()
                        if full:
                            return solver.binary_resume(dCs, dCt, dA, shape, st, 20000, eb,
                                                        cc, steps, stream=stream)
                        return solver.binary(dCs, dCt, dA, shape, 20000, eb, cc, steps,
                                             u0=st["u"], pt0=st["pt"], stream=stream)

                    tq, rq, _ = _median_call(warm_eb)
                    eq = solver.energy_binary(dCs, dCt, dA, rq.u, shape, stream=stream)[0]
                    gq = (eq - e_star) / abs(e_star)
                    if gq <= gap_cold:
                        out[key]["same_energy"] = {
                            "err_bound": eb, "iterations": rq.iterations, "seconds": tq,
                            "energy_gap": gq,
                            "speedup_vs_cold": out["time_to_tol"]["seconds"] / tq}
                        break
                del net, st, rw
            # the bound on any full-state predictor: resume from the reference's own state
            rs = solver.binary(dCs, dCt, dA, shape, REF_ITERS, 0.0, cc, steps, stream=stream,
                               want_state=True)
            ideal = {"u": rs.u, **rs.state}
            def resume_ideal():
                return solver.binary_resume(dCs, dCt, dA, shape, ideal, 20000, 1e-4, cc,
                                            steps, stream=stream)

            resume_ideal()
            t_ideal, ri, _ = _median_call(resume_ideal)
            if "warm_start" in out:
                out["warm_start"]["ideal_full_state"] = {
                    "seconds": t_ideal, "iterations": ri.iterations,
                    "note": "resume from the 2000-iteration solve's own state: what an exact "
                            "full-state predictor would reach, inference excluded"}
            del rs, ideal, ri

    _progress("GPU legs done")
    if not args.no_cpu:
        orc = _oracle()
        k = max(1, args.cpu_iters)
        cpu_s, runs = _cpu_time(lambda: orc.binary(Cs, Ct, alpha, shape, k, 0.0, cc, steps),
                                args.cpu_repeats)
        model, ncpu = _host_cpu()
        rate = N * k / cpu_s
        out["cpu_baseline"] = {"value": rate, "unit": "voxel-iter/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations of the same {n}^3 volume "
                                         f"(FP32 reference-order oracle, gcc -O2, 1 thread "
                                         f"pinned to one core), median of "
                                         f"{len(runs)} runs of "
                                         f"{', '.join(f'{t:.2f}' for t in runs)} s incl. init",
                               "host_cpu": model, "host_logical_cpus": ncpu}
        if "time_to_tol" in out:
            it = out["time_to_tol"]["iterations"]
            out["cpu_baseline"]["time_to_tol_extrapolated_s"] = N * it / rate
    return out


def bench_binary2d(args):
    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    Cs, Ct, a, pars, _ = workloads.binary2d_ellipses(256, seed=1001)
    shape, N, iters = (256, 256), 256 * 256, 300
    d = [solver.to_device(x) for x in (Cs, Ct, a)]
    for _ in range(args.warmup):
        solver.binary(*d, shape, iters, 0.0, 0.3, 0.16)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        r = solver.binary(*d, shape, iters, 0.0, 0.3, 0.16, time_kernel=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    out = _base(args, "voxel-iters/s, 256^2 binary MF 2D (300 fixed iterations)",
                N * iters * args.steps / wall, "voxel-iter/s", wall * 1e3 / args.steps, "f32",
                "synthetic (workloads.binary2d_ellipses seed 1001)",
                {"workload": "binary2d 256^2 ellipses, alpha 0.5, cc 0.3 steps 0.16",
                 "iters_per_step": iters})
    out["roofline"] = _roofline(52, N, r.kernel_ms)
    out["roofline"]["note"] = "3.4 MB per iteration: cache-resident and launch-bound"
    if not args.no_cpu:
        orc = _oracle()
        t = time.perf_counter()
        ref = orc.binary(Cs, Ct, a, shape, iters, 0.0, 0.3, 0.16)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": N * iters / cpu_s, "unit": "voxel-iter/s", "cores": 1,
                               "kind": "port",
                               "sample": f"the full 300-iteration solve, {cpu_s:.1f} s"}
        out["u_bit_identical_to_cpu"] = bool(np.array_equal(
            r.u.cpu().numpy().view(np.uint32), ref.u.view(np.uint32)))
    return out


def bench_potts3d(args):
    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    shape, L, iters = (192, 192, 192), 6, 300
    Ct, a, _ = workloads.potts3d_voronoi(shape, nlab=L, seed=1003)
    N = int(np.prod(shape))
    d = [solver.to_device(x) for x in (Ct, a)]
    cc, st = workloads.POTTS3D_CC, workloads.POTTS3D_STEPS
    kern = "passes" if args.kernel == "passes" else "fused"
    for _ in range(args.warmup):
        solver.potts(*d, shape, L, 5, 0.0, cc, st, kernel=kern, tile_x=args.tile_x,
                     zchunk=args.zchunk)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        r = solver.potts(*d, shape, L, iters, 0.0, cc, st, kernel=kern, tile_x=args.tile_x,
                         zchunk=args.zchunk, time_kernel=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    out = _base(args, "voxel-iters/s, 192^3 Potts L=6 (300 fixed iterations)",
                N * iters * args.steps / wall, "voxel-iter/s", wall * 1e3 / args.steps, "f32",
                "synthetic (workloads.potts3d_voronoi seed 1003)",
                {"workload": "potts3d 192^3 L=6 Voronoi, alpha 0.8, cc 0.25 steps 0.11",
                 "iters_per_step": iters, "kernel": kern})
    out["roofline"] = _roofline(48 * L + 8, N, r.kernel_ms)
    lab = r.u.view(L, -1).argmax(0)
    out["label_fractions"] = [float((lab == l).float().mean()) for l in range(L)]
    if not args.no_cpu:
        orc = _oracle()
        k = max(1, args.cpu_iters // 4)
        t = time.perf_counter()
        orc.potts(Ct, a, shape, L, k, 0.0, cc, st)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": N * k / cpu_s, "unit": "voxel-iter/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations incl. init, FP32 reference-order "
                                         f"oracle, {cpu_s:.1f} s"}
    return out


def bench_ishikawa3d(args):
    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    shape, L, iters = (192, 192, 192), 6, 300
    Ct, a, _ = workloads.ishikawa3d_voronoi(shape, nlab=L, seed=1003)
    N = int(np.prod(shape))
    d = [solver.to_device(x) for x in (Ct, a)]
    cc, st = workloads.ISHIKAWA3D_CC, workloads.ISHIKAWA3D_STEPS
    for _ in range(args.warmup):
        solver.ishikawa(*d, shape, L, 5, 0.0, cc, st, tile_x=args.tile_x, zchunk=args.zchunk)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        r = solver.ishikawa(*d, shape, L, iters, 0.0, cc, st, tile_x=args.tile_x,
                            zchunk=args.zchunk, time_kernel=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    out = _base(args, "voxel-iters/s, 192^3 Ishikawa L=6 (300 fixed iterations)",
                N * iters * args.steps / wall, "voxel-iter/s", wall * 1e3 / args.steps, "f32",
                "synthetic (workloads.ishikawa3d_voronoi seed 1003)",
                {"workload": "ishikawa3d 192^3 L=6 ordered Voronoi labels, alpha 0.3/(l+1), "
                             "cc 0.75 steps 0.1", "iters_per_step": iters})
    out["roofline"] = _roofline(12 * L + 36 * (L - 1), N, r.kernel_ms)
    lab = (r.u.view(L - 1, -1) > 0.5).sum(0)
    out["label_fractions"] = [float((lab == l).float().mean()) for l in range(L)]
    if not args.no_cpu:
        orc = _oracle()
        k = max(1, args.cpu_iters // 4)
        t = time.perf_counter()
        orc.ishikawa(Ct, a, shape, L, k, 0.0, cc, st)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": N * k / cpu_s, "unit": "voxel-iter/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations incl. init, FP32 reference-order "
                                         f"oracle, {cpu_s:.1f} s"}
    return out


def _hmf_bytes(parent, nd):
    """Algorithmic bytes per voxel-iteration of the fused HMF path: sink (read div, u and
    Ct or the previous pt, write pt: 16 B per non-root node, 8 B at the root) plus the
    fused flow kernel per non-root node (read u, pt, parent pt and nd faces, write u, nd
    faces and div: 44 B in 3D, 36 B in 2D; scalar alpha)."""
    nflow = sum(1 for p in parent if p >= 0)
    return 8 + nflow * (16 + (44 if nd == 3 else 36))


def _bench_hmf(args, parent, Ct, alpha, shape, iters, steps, cc, metric, data, workload):
    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import hmf

    N = int(np.prod(shape))
    kern = "passes" if args.kernel == "passes" else "fused"
    for _ in range(args.warmup):
        hmf.solve_tree(parent, Ct, alpha, shape, 3, steps, cc, kernel=kern)
    torch.cuda.synchronize()
    # inputs resident on the device before the timed region
    dev = [None if c is None else hmf._flat_dev(c, shape).view(*shape[::-1]).permute(
        *range(len(shape) - 1, -1, -1)) for c in Ct]
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        r = hmf.solve_tree(parent, dev, alpha, shape, iters, steps, cc, kernel=kern,
                           timing=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    out = _base(args, metric, N * iters * args.steps / wall, "voxel-iter/s",
                wall * 1e3 / args.steps, "f32", data,
                {"workload": workload, "nodes": len(parent), "iters_per_step": iters,
                 "kernel": kern})
    out["roofline"] = _roofline(_hmf_bytes(parent, len(shape)), N, r.kernel_ms)
    out["data_cost"] = float(r.data_cost[0])
    out["regu_cost"] = float(r.regu_cost[0])
    if not args.no_cpu:
        orc = _oracle()
        k = max(1, args.cpu_iters // 10)
        t = time.perf_counter()
        orc.hmf(parent, Ct, alpha, shape, k, steps, cc)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": N * k / cpu_s, "unit": "voxel-iter/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{k} iterations incl. init, FP32 reference-order "
                                         f"oracle, {cpu_s:.1f} s"}
    return out


def bench_hmf2d(args):
    from asetsmatlabmaxflow_amd import workloads as W

    L = 40
    costs, _, _ = W.hmf_linear2d((256, 256), L, seed=1005)
    parent = [-1] + [0] * L
    return _bench_hmf(args, parent, [None] + costs, [0.0] + [W.HMF2D_ALPHA] * L, (256, 256),
                      W.HMF2D_ITERS, W.HMF2D_STEPS, 1.0 / L,
                      "voxel-iters/s, 256^2 HMF linear model, 40 leaves (200 fixed iterations)",
                      "synthetic (workloads.hmf_linear2d seed 1005)",
                      "hmf2d 256^2 flat tree of 40 leaves, Ct_i = sqrt(|img - i/39|), alpha "
                      "0.15, cc 1/40 steps 0.1 (runTestHMFLinearModel.m)")


def bench_hmf3d(args):
    from asetsmatlabmaxflow_amd import workloads as W

    shape = (192, 192, 192)
    parent, Ct, alpha, _ = W.hmf3d_tree(shape, seed=1006)
    return _bench_hmf(args, parent, Ct, alpha, shape, W.HMF3D_ITERS, W.HMF3D_STEPS, W.HMF3D_CC,
                      "voxel-iters/s, 192^3 HMF two-level tree, 6 leaves (100 fixed iterations)",
                      "synthetic (workloads.hmf3d_tree seed 1006)",
                      "hmf3d 192^3 root -> 2 branches -> 3 leaves each, Voronoi intensities, "
                      "alpha 0.5 branches 1.0 leaves, cc 0.25 steps 0.11")


def bench_batched(args):
    import numpy as np
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    rng = np.random.default_rng(1004)
    sizes = [tuple(int(v) for v in rng.integers(64, 193, 3)) for _ in range(args.nvol)]
    dev = []
    t = time.perf_counter()
    for v, sh in enumerate(sizes):
        Cs, Ct, a, _ = workloads.binary3d_ellipsoids(sh, seed=1004 * 1000 + v)
        dev.append(tuple(solver.to_device(x) for x in (Cs, Ct, a)) + (sh,))
    gen_s = time.perf_counter() - t
    cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    res, legs = {}, {}
    for flow in ("f32", "f16"):
        solver.binary_batched(dev, 3, 0.0, cc, st, flow=flow)   # autotuning + warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        rs, secs, loop_ms = solver.binary_batched(dev, 2000, 1e-4, cc, st, flow=flow,
                                                  time_kernel=True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        vit = sum(int(np.prod(s)) * r.iterations for s, r in zip(sizes, rs))
        bpv = 60 if flow == "f32" else 40
        res[flow] = rs
        legs[flow] = {"seconds": wall, "loop_ms": loop_ms, "voxel_iterations": vit,
                      "voxel_iter_per_s": vit / wall,
                      "achieved_GBps_loop": bpv * vit / (loop_ms * 1e-3) / 1e9,
                      "iterations": [r.iterations for r in rs]}
    diff = max(float((a.u - b.u).abs().max().item()) for a, b in zip(res["f32"], res["f16"]))
    agree = np.mean([float(((a.u > 0.5) == (b.u > 0.5)).float().mean().item())
                     for a, b in zip(res["f32"], res["f16"])])
    out = _base(args, f"voxel-iters/s, {args.nvol} uneven binary volumes batched to errBound "
                      "1e-4 (fp32 flows)", legs["f32"]["voxel_iter_per_s"], "voxel-iter/s",
                legs["f32"]["seconds"] * 1e3, "f32",
                "synthetic (workloads.binary3d_ellipsoids, seeds 1004*1000+v)",
                {"workload": f"batched binary3d, {args.nvol} volumes, axes U{{64..192}}",
                 "total_voxels": int(sum(np.prod(s) for s in sizes))})
    out["roofline"] = {"bound": "hbm", "achieved": legs["f32"]["achieved_GBps_loop"],
                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": legs["f32"]["achieved_GBps_loop"] / HBM_PEAK_GBS,
                       "traffic": None}
    out["fp32"] = legs["f32"]
    out["fp16"] = legs["f16"]
    out["fp16_vs_fp32"] = {"max_abs_u_diff": diff, "label_agreement": agree}
    out["gen_s"] = gen_s
    if not args.no_cpu:
        # CPU baseline: the FP32 reference-order oracle on the volumes in index order
        # until about 10 s of one pinned core have been spent, each run for exactly the
        # iterations its GPU solve executed (and its u compared bit for bit); the FP64
        # oracle on the three smallest volumes is the error yardstick of both storages
        orc = _oracle()
        cpu_s, vit_cpu, nv, bitexact = 0.0, 0, 0, True
        for v, sh in enumerate(sizes):
            if cpu_s >= 10.0:
                break
            Cs, Ct, a, _ = workloads.binary3d_ellipsoids(sh, seed=1004 * 1000 + v)
            it32 = res["f32"][v].iterations
            box = {}
            t1, _ = _cpu_time(lambda: box.update(
                r=orc.binary(Cs, Ct, a, sh, it32, 0.0, cc, st)), 1)
            cpu_s += t1
            vit_cpu += int(np.prod(sh)) * it32
            nv += 1
            bitexact &= bool(np.array_equal(res["f32"][v].u.cpu().numpy().view(np.uint32),
                                            box["r"].u.view(np.uint32)))
        model, ncpu = _host_cpu()
        out["cpu_baseline"] = {
            "value": vit_cpu / cpu_s, "unit": "voxel-iter/s", "cores": 1, "kind": "port",
            "sample": f"volumes 0..{nv - 1} for the iterations their GPU solves executed, "
                      f"FP32 reference-order oracle, 1 pinned core, {cpu_s:.1f} s",
            "host_cpu": model, "host_logical_cpus": ncpu,
            "u_fp32_bit_identical_to_cpu": bitexact}
        pick = sorted(range(len(sizes)), key=lambda v: int(np.prod(sizes[v])))[:3]
        err = {"fp32": [0.0, 0.0, 1.0], "fp16": [0.0, 0.0, 1.0]}   # max, mean, labels
        for v in pick:
            Cs, Ct, a, _ = workloads.binary3d_ellipsoids(sizes[v], seed=1004 * 1000 + v)
            for flow, key in (("f32", "fp32"), ("f16", "fp16")):
                r64 = orc.binary(Cs, Ct, a, sizes[v], res[flow][v].iterations, 0.0, cc, st,
                                 precision="f64")
                ug = res[flow][v].u.cpu().numpy().astype(np.float64)
                d = np.abs(ug - r64.u)
                e = err[key]
                err[key] = [max(e[0], float(d.max())), max(e[1], float(d.mean())),
                            min(e[2], float(np.mean((ug > 0.5) == (r64.u > 0.5))))]
        out["error_vs_fp64_oracle"] = {
            key: {"max_abs_u_err": e[0], "mean_abs_u_err": e[1], "label_agreement": e[2]}
            for key, e in err.items()}
        out["error_vs_fp64_oracle"]["volumes"] = ["x".join(map(str, sizes[v])) for v in pick]
        out["error_vs_fp64_oracle"]["note"] = (
            "worst volume of the sample per figure; each GPU result against the FP64 "
            "oracle run for the same iteration count")
    return out


def bench_hbm(args):
    """HBM ceiling: plain 16-byte-per-lane streams (read, copy, and the binary solver's
    traffic mix without its stencil) over about 1 GB, each swept over grid size, cache
    policy and loads in flight; the best of each is the ceiling that kernel shape gets."""
    import torch

    from asetsmatlabmaxflow_amd import solver

    dev = torch.device("cuda:0")
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    N = 256 ** 3
    buf = torch.ones(60 * N // 4, dtype=torch.float32, device=dev)   # 1006.6 MB
    torch.cuda.synchronize()
    rows = []
    for mode in ("read", "copy", "mix", "mix_packed", "mix_state", "mix_blk"):
        for nt in (True, False):
            for one in ((False, True) if mode in ("read", "copy") else (False,)):
                for per_cu in (2, 4, 8, 16):
                    ms, traffic = solver.hbm_stream(mode, buf, nontemporal=nt,
                                                    one_in_flight=one, grid=per_cu * cus,
                                                    iters=args.steps)
                    rows.append({"mode": mode, "nontemporal": nt,
                                 "in_flight_per_lane": 1 if one else 4,
                                 "workgroups_per_cu": per_cu, "ms": ms, "bytes": traffic,
                                 "GBps": traffic / (ms * 1e-3) / 1e9})
    best = {m: max((r for r in rows if r["mode"] == m), key=lambda r: r["GBps"])
            for m in ("read", "copy", "mix", "mix_packed", "mix_state", "mix_blk")}
    mix = best["mix"]
    out = _base(args, "HBM GB/s ceiling: plain streams with 16-byte accesses (read, copy, "
                      "binary-solver traffic mix)", mix["GBps"], "GB/s", mix["ms"], "f32",
                "synthetic (ones)",
                {"workload": "hbm streams over 1006.6 MB (mix: 256^3 voxels x 60 B)",
                 "cus": cus})
    out["roofline"] = {"bound": "hbm", "achieved": mix["GBps"], "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": mix["GBps"] / HBM_PEAK_GBS, "traffic": mix["bytes"],
                       "traffic_source": "bytes the stream moves by construction"}
    out["best"] = best
    out["sweep"] = rows
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="binary3d",
                    choices=["binary3d", "binary2d", "potts3d", "batched", "ishikawa3d", "hmf2d",
                             "hmf3d", "hbm"])
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--nvol", type=int, default=64)
    ap.add_argument("--kernel", default="fused", choices=["fused", "passes"])
    ap.add_argument("--tile-x", type=int, default=0)
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--cpu-iters", type=int, default=int(os.environ.get("ASETS_CPU_ITERS", 20)))
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--no-tol", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the workload, run three device-to-device copies of MIB MiB "
                         "(a kernel of known traffic to scale rocprofv3 byte counters by)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    fn = {"binary3d": bench_binary3d, "binary2d": bench_binary2d, "potts3d": bench_potts3d,
          "batched": bench_batched, "ishikawa3d": bench_ishikawa3d, "hmf2d": bench_hmf2d,
          "hmf3d": bench_hmf3d, "hbm": bench_hbm}[args.workload]
    out = fn(args)
    if args.copy_calibration > 0:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        out["copy_calibration_mib"] = args.copy_calibration
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
