"""Calibration of per-road speed multipliers and the soft-avoidance strength
by gradient descent through the differentiable tick.

A hidden-parameter run (theta_true, alpha_true) of the scenarios gives the
observed final positions; starting from theta = 1 and alpha = alpha0, Adam
minimises the mean squared distance between simulated and observed final
positions, with gradients from the hand-written adjoint (msim.adjoint).  By
default the adjoint differentiates arrivals as carrying their leftover time
onto the next road (arrival_gradient="carry"): with the frozen run's own
gradient ("frozen"), an arrival's snap cuts every earlier road's influence
and the multipliers barely improve.
Scenarios share one map, so theta has one entry per map road and is tied
across the scenarios' packed copies.  theta = exp(phi) and alpha = exp(psi)
keep both positive.

    python -m msim.calibrate [--scenarios 64] [--agents 100000] [--ticks 1000] [--steps 50]
                             [--arrival-gradient frozen|carry]
"""
import argparse
import json
import time

import numpy as np
import torch

from . import scenarios as S
from ._lib import Batch
from .adjoint import DiffRun


def calibrate(scs, ticks=1000, steps=50, lr=0.01, lr_alpha=0.05, alpha_true=0.02, alpha0=0.01, sigma=0.2,
              checkpoint_every=50, seed=0, arrival_gradient="carry", log=None):
    dev = torch.device("cuda")
    n_map = scs[0].roads.shape[0]
    rng = np.random.default_rng(seed)
    theta_true = torch.from_numpy(np.exp(rng.normal(0.0, sigma, n_map)).astype(np.float32)).to(dev)
    reps = len(scs)

    truth = DiffRun(Batch(scs), ticks, checkpoint_every)
    t0 = time.perf_counter()
    observed = truth(theta_true.repeat(reps), torch.tensor(alpha_true, device=dev)).detach()
    torch.cuda.synchronize()
    t_hidden = time.perf_counter() - t0
    truth.batch.close()

    run = DiffRun(Batch(scs), ticks, checkpoint_every, arrival_gradient=arrival_gradient)
    phi = torch.zeros(n_map, device=dev, requires_grad=True)
    psi = torch.tensor(float(np.log(alpha0)), device=dev, requires_grad=True)
    opt = torch.optim.Adam([{"params": [phi], "lr": lr}, {"params": [psi], "lr": lr_alpha}])
    hist = []
    grad_s = []
    for step in range(steps + 1):
        opt.zero_grad()
        t0 = time.perf_counter()
        alpha = torch.exp(psi)
        pos = run(torch.exp(phi).repeat(reps), alpha)
        loss = ((pos - observed) ** 2).sum(dim=1).mean()
        if step < steps:
            loss.backward()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        with torch.no_grad():
            err = float(((torch.exp(phi) - theta_true).abs() / theta_true).mean())
        hist.append({"step": step, "loss_m2": float(loss.detach()), "alpha": float(alpha.detach()),
                     "theta_mean_rel_err": err, "seconds": dt})
        if log:
            log(f"step {step}: loss {float(loss.detach()):.4f} m^2, alpha {float(alpha.detach()):.5f}, "
                f"mean |theta - theta_true| / theta_true {err:.4f}, {dt:.2f} s")
        if step < steps:
            grad_s.append(dt)
            opt.step()
    run.batch.close()
    n = int(sum(sc.entities.shape[0] for sc in scs))
    return {
        "agents": n,
        "roads_per_map": int(n_map),
        "ticks": ticks,
        "steps": steps,
        "arrival_gradient": arrival_gradient,
        "alpha_true": alpha_true,
        "hidden_run_seconds": t_hidden,
        "seconds_per_gradient": float(np.median(grad_s)) if grad_s else None,
        "best_step": int(np.argmin([h["loss_m2"] for h in hist])),
        "history": hist,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenarios", type=int, default=64)
    ap.add_argument("--agents", type=int, default=100_000)
    ap.add_argument("--nodes", type=int, default=20_000)
    ap.add_argument("--ticks", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--lr", type=float, default=0.01)
    ap.add_argument("--arrival-gradient", choices=("frozen", "carry"), default="carry")
    a = ap.parse_args()
    scs = S.adjoint_config(a.scenarios, a.agents, a.nodes)
    out = calibrate(scs, a.ticks, a.steps, lr=a.lr, arrival_gradient=a.arrival_gradient,
                    log=lambda m: print(m, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
