"""Train the U-Net pressure guess on pairs solved by FP64 conjugate gradients.

    python -m natrix_amd.learned.train --samples 20480 --size 128 --epochs 10 \\
        --out natrix_amd/learned/unet_weights.pt

Loss: mean squared error of the mean-free prediction against the mean-free
target pressure over fluid cells (the operator fixes p only up to a
constant), relative to the target's mean square.  The problem is invariant
under the 8 symmetries of the square, so every batch is randomly flipped and
transposed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from .unet import UNet, pack_input

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "unet_weights.pt")


def masked_loss(pred, target, fluid):
    f = fluid.float()
    cnt = f.sum(dim=(1, 2)).clamp_min(1.0)
    pm = (pred * f).sum(dim=(1, 2)) / cnt
    tm = (target * f).sum(dim=(1, 2)) / cnt
    d = ((pred - pm[:, None, None]) - (target - tm[:, None, None])) * f
    t = (target - tm[:, None, None]) * f
    return ((d * d).sum(dim=(1, 2)) / (t * t).sum(dim=(1, 2)).clamp_min(1e-12)).mean()


def augment(x, *ys, gen):
    k = int(torch.randint(0, 8, (1,), generator=gen))
    def t(a, nd):
        # spatial dims are 1 and 2 for [N,H,W] and [N,H,W,C]
        if k & 1:
            a = a.flip(1)
        if k & 2:
            a = a.flip(2)
        if k & 4:
            a = a.transpose(1, 2)
        return a.contiguous()
    return (t(x, 4),) + tuple(t(y, 3) for y in ys)


def train(b, solid, p, epochs=6, batch=32, lr=1e-3, seed=0, log=print, val_frac=0.05,
          time_budget=None):
    dev = "cuda"
    torch.manual_seed(seed)
    gen = torch.Generator().manual_seed(seed)
    n = len(b)
    nval = max(1, int(n * val_frac))
    B = torch.from_numpy(b).to(dev)
    S = torch.from_numpy(solid).to(dev)
    P = torch.from_numpy(p).to(dev)
    net = UNet().to(dev)
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    steps = epochs * ((n - nval) // batch)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=max(1, steps))
    hist = []
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(n - nval, generator=gen) + nval
        tot = 0.0
        nb = 0
        for i in range(0, n - nval - batch + 1, batch):
            idx = perm[i:i + batch].to(dev)
            x = pack_input(B[idx], S[idx])
            x, tgt, fl = augment(x, P[idx], ~S[idx], gen=gen)
            out = net(x)
            loss = masked_loss(out, tgt, fl)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            sched.step()
            tot += float(loss)
            nb += 1
            if nb % 100 == 0:
                log(json.dumps({"epoch": ep, "batch": nb, "loss": tot / nb,
                                "seconds": round(time.time() - t0, 1)}))
            if time_budget is not None and time.time() - t0 > time_budget:
                break
        with torch.no_grad():
            vl = []
            for i in range(0, nval, batch):
                sl = slice(i, min(nval, i + batch))
                vl.append(float(masked_loss(net(pack_input(B[sl], S[sl])), P[sl], ~S[sl])))
        rec = {"epoch": ep, "train_loss": tot / max(nb, 1), "val_loss": float(np.mean(vl)),
               "seconds": round(time.time() - t0, 1)}
        hist.append(rec)
        log(json.dumps(rec))
        if time_budget is not None and time.time() - t0 > time_budget:
            log(json.dumps({"stopped": "time budget", "seconds": round(time.time() - t0, 1)}))
            break
    return net, hist


def save(net, path, meta):
    sd = {k: v.detach().to(torch.bfloat16).cpu() for k, v in net.state_dict().items()}
    torch.save({"state": sd, "meta": meta}, path)


def load(path=DEFAULT_WEIGHTS, device="cuda"):
    blob = torch.load(path, map_location="cpu", weights_only=True)
    net = UNet()
    net.load_state_dict({k: v.float() for k, v in blob["state"].items()})
    return net.to(device).eval(), blob.get("meta", {})


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=20000,
                    help="FP64 conjugate-gradient iterations per training problem")
    ap.add_argument("--targets", choices=["fp64", "jacobi"], default="fp64",
                    help="fp64: FP64 conjugate-gradient solutions to --tol; jacobi: the "
                         "batched float32 Jacobi solver run to its round-off floor")
    ap.add_argument("--seed", type=int, default=0, help="initialisation and shuffling seed")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--time-budget", type=float, default=None,
                    help="stop training after this many seconds (weights are saved)")
    a = ap.parse_args(argv)
    from .data import make_dataset, make_dataset_jacobi
    log = lambda m: print(m, flush=True)
    t0 = time.time()
    if a.targets == "fp64":
        b, solid, p, res = make_dataset(a.samples, a.size, tol=a.tol, max_iter=a.max_iter,
                                        log=log)
    else:
        b, solid, p, res = make_dataset_jacobi(a.samples, a.size, tol=a.tol, log=log)
    # a target the solver did not bring to the tolerance is not a solution:
    # such samples are left out (and counted)
    keep = res <= 10.0 * a.tol if a.targets == "fp64" else np.isfinite(res)
    dropped = int((~keep).sum())
    if dropped:
        b, solid, p = b[keep], solid[keep], p[keep]
    edges = [0.0] + [10.0 ** e for e in range(-12, 0)] + [float("inf")]
    hist, _ = np.histogram(res, bins=edges)
    residual_hist = {f"<={edges[i + 1]:g}": int(c) for i, c in enumerate(hist) if c}
    log(json.dumps({"dataset": a.samples, "size": a.size, "seconds": round(time.time() - t0, 1),
                    "residual_median": float(np.median(res)), "residual_max": float(res.max()),
                    "frac_at_tol": float((res <= a.tol).mean()),
                    "dropped_unconverged": dropped, "residual_histogram": residual_hist}))
    net, hist = train(b, solid, p, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                      log=log, time_budget=a.time_budget)
    meta = {"samples": a.samples, "size": a.size, "epochs": a.epochs,
            "data": ("FP64 conjugate gradients (nx_fp64_solve_batch), float32 rounding of "
                     "the FP64 solution" if a.targets == "fp64" else
                     "batched float32 Jacobi from zero to its round-off floor"),
            "seed": a.seed,
            "data_tol": a.tol, "data_residual_median": float(np.median(res)),
            "data_residual_max": float(res.max()), "data_residual_histogram": residual_hist,
            "data_dropped_unconverged": dropped,
            "history": hist, "params": net.n_params()}
    save(net, a.out, meta)
    log(json.dumps({"saved": a.out, "params": net.n_params()}))


if __name__ == "__main__":
    main(sys.argv[1:])
