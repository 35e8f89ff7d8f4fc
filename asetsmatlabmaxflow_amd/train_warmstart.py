"""Train the warm-start network on synthetic binary volumes solved by the plain GPU solver.

    python -m asetsmatlabmaxflow_amd.train_warmstart --volumes 256 --size 128 --steps 2000 \
        --batch 4 --out weights.npz

A long run can be split: --checkpoint saves weights, AdamW moments and the step count,
--resume continues from them, and --total-steps keeps one cosine schedule across parts.

Data: volumes of the 256^3 ellipsoid generator (workloads.binary3d_ellipsoids) at
--size^3, seeds 2000, 2001, ...; targets are the plain solver's state after
--target-iters fixed iterations (errBound 0): (u, pt) for --variant full, with loss
mean (u0-u)^2 + mean ((pt0-pt)/10)^2; (u, bx, by, bz) for --variant s2d, with loss
mean (u0-u)^2 + 2 * mean of the squared face-flow errors.
Optimiser: AdamW (lr --lr, betas 0.9/0.999, weight decay 0.01), cosine decay to 10 %.
Every forward/backward/optimiser operation is a kernel of libasets_nn.so.
Prints one progress line every --log-every steps and a JSON summary at the end.
"""
from __future__ import annotations

import argparse
import json
import math
import multiprocessing as mp
import os
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np


def _gen(args):
    size, seed = args
    from asetsmatlabmaxflow_amd import workloads

    Cs, Ct, a, _ = workloads.binary3d_ellipsoids((size, size, size), seed=seed)
    return tuple(np.asarray(x, np.float32).ravel(order="F") for x in (Cs, Ct, a))


def make_dataset(nvol: int, size: int, seed0: int, target_iters: int, device, workers: int,
                 log=print, flows: bool = False):
    """Generate nvol volumes on the CPU (process pool), solve each on the GPU, and return
    device tensors Cs, Ct, alpha, tu, tpt (and with flows the face flows tbx, tby, tbz)
    of shape [nvol * size^3]."""
    import torch

    from asetsmatlabmaxflow_amd import solver, workloads

    N = size ** 3
    shape = (size, size, size)
    keys = ("Cs", "Ct", "alpha", "tu") + (("tbx", "tby", "tbz") if flows else ("tpt",))
    bufs = {k: torch.empty(nvol * N, dtype=torch.float32, device=device) for k in keys}
    cc, st = workloads.BINARY3D_CC, workloads.BINARY3D_STEPS
    t0 = time.time()
    # spawned workers (numpy only) never inherit this process's GPU context
    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        for v, (Cs, Ct, a) in enumerate(ex.map(_gen, [(size, seed0 + i) for i in range(nvol)])):
            sl = slice(v * N, (v + 1) * N)
            for k, arr in (("Cs", Cs), ("Ct", Ct), ("alpha", a)):
                bufs[k][sl].copy_(torch.from_numpy(arr))
            r = solver.binary(bufs["Cs"][sl], bufs["Ct"][sl], bufs["alpha"][sl], shape,
                              target_iters, 0.0, cc, st, want_state=True, out=bufs["tu"][sl])
            if flows:
                for k in ("bx", "by", "bz"):
                    bufs["t" + k][sl].copy_(r.state[k])
            else:
                bufs["tpt"][sl].copy_(r.state["pt"])
            del r
            if (v + 1) % max(1, nvol // 8) == 0:
                log(f"dataset {v + 1}/{nvol} volumes, {time.time() - t0:.1f} s")
    torch.cuda.synchronize()
    return bufs


def train(net, data, size: int, nvol: int, steps: int, batch: int, lr: float, seed: int,
          log_every: int = 50, log=print, total_steps: int | None = None,
          flow_weight: float | None = None):
    """Run `steps` AdamW steps from the network's own step count (0, or where a
    checkpoint left it) on a cosine schedule over `total_steps` (default: steps).  Each
    step takes `batch` volumes drawn at random (batch groups of consecutive volumes)."""
    import torch

    N = size ** 3
    shape = (size, size, size)
    ngroups = nvol // batch
    step0 = net.step_count
    total = total_steps or (step0 + steps)
    rng = np.random.default_rng([seed, step0])
    hist = []
    t0 = time.time()
    for i in range(steps):
        s = step0 + i
        g = int(rng.integers(ngroups))
        sl = slice(g * batch * N, (g + 1) * batch * N)
        Cs, Ct, a = data["Cs"][sl], data["Ct"][sl], data["alpha"][sl]
        if "tbx" in data:   # space-to-depth network, full state
            pred = net.forward(Cs, Ct, a, shape, batch, train=True)
            kw = {} if flow_weight is None else {"lam": flow_weight}
            loss = net.backward_state(pred, {k: data["t" + k][sl] for k in
                                             ("u", "bx", "by", "bz")}, shape, batch, **kw)
        else:
            u0, pt0 = net.forward(Cs, Ct, a, shape, batch, train=True)
            loss = net.backward(Cs, Ct, u0, pt0, data["tu"][sl], data["tpt"][sl], shape,
                                batch)
        lr_s = lr * (0.1 + 0.9 * 0.5 * (1 + math.cos(math.pi * s / max(1, total))))
        net.adamw_step(lr=lr_s)
        if (i + 1) % log_every == 0 or i == 0:
            lv = float(loss)
            hist.append((s + 1, lv))
            log(f"step {s + 1}/{total} loss {lv:.5f} lr {lr_s:.2e} "
                f"{(time.time() - t0) / (i + 1) * 1e3:.1f} ms/step")
    torch.cuda.synchronize()
    return hist, time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--volumes", type=int, default=64)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--seed0", type=int, default=2000)
    ap.add_argument("--target-iters", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--log-every", type=int, default=50)
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    ap.add_argument("--resume", default=None,
                    help="continue from this checkpoint (weights, AdamW moments, step count)")
    ap.add_argument("--total-steps", type=int, default=0,
                    help="length of the cosine schedule (default: this run's last step)")
    ap.add_argument("--checkpoint", default=None, help="write a checkpoint here at the end")
    ap.add_argument("--variant", default="full", choices=["full", "s2d"],
                    help="full: WarmStartNet predicting (u0, pt0); s2d: WarmStartNetS2D "
                         "predicting (u0, bx0, by0, bz0)")
    ap.add_argument("--flow-weight", type=float, default=None,
                    help="s2d: weight of the face-flow errors in the loss (default 2)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    import torch

    from asetsmatlabmaxflow_amd.warmstart import WarmStartNet, WarmStartNetS2D

    dev = torch.device("cuda:0")
    t0 = time.time()
    data = make_dataset(args.volumes, args.size, args.seed0, args.target_iters, dev,
                        args.workers, log=lambda m: print(m, flush=True),
                        flows=args.variant == "s2d")
    data_s = time.time() - t0
    net = (WarmStartNetS2D if args.variant == "s2d" else WarmStartNet)(dev, seed=args.seed)
    if args.init:
        net.load(args.init)
    if args.resume:
        net.load_checkpoint(args.resume)
    hist, train_s = train(net, data, args.size, args.volumes, args.steps, args.batch, args.lr,
                          args.seed, args.log_every, log=lambda m: print(m, flush=True),
                          total_steps=args.total_steps or None, flow_weight=args.flow_weight)
    net.save(args.out)
    if args.checkpoint:
        net.save_checkpoint(args.checkpoint)
    print(json.dumps({"variant": args.variant, "volumes": args.volumes, "size": args.size,
                      "steps": args.steps,
                      "batch": args.batch, "lr": args.lr, "target_iters": args.target_iters,
                      "dataset_s": data_s, "train_s": train_s,
                      "ms_per_step": train_s / max(1, args.steps) * 1e3,
                      "loss_first": hist[0][1], "loss_last": hist[-1][1], "history": hist,
                      "weights": args.out}), flush=True)


if __name__ == "__main__":
    main()
