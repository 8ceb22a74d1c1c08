"""Train and evaluate the learned preconditioner on the GPU.

Data come only from the plain solver: for every training seed the κ field is
assembled and solved with Jacobi-PCG to 1e-8 (the "solution"), and Jacobi-PCG
iterates ``x_k`` at a few iteration counts ``k`` give residuals
``r_k = b - A x_k`` and errors ``e_k = x* - x_k``.  The loss for one residual
is the relative energy error of the network's correction

    E = ||y - e~||_A^2 / ||e~||_A^2 = (y.A^y - 2 y.r~) / (e~.r~) + 1

(normalised system, ``e~ = D^1/2 e / s``); its gradient only needs ``A^ y - r~``
so the solution enters as the normaliser ``e~.r~ = e.r / s^2``.  Each step
re-assembles the sample's matrix on the device, refreshes the level values,
and accumulates the gradients of all its residuals before one Adam step.

    python -m tensormesh_amd.precond.train --n 96 --train 2000 --out weights.npz
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time
import warnings
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .. import _lib
from ..sparse.ops import _spmv_raw, pcg_raw
from .fcg import LearnedFCG
from .network import Adam, GNNPreconditioner
from .problem import PoissonCube


def log(msg: str):
    print(f"[precond] {msg}", flush=True)


def plain_solve(prob: PoissonCube, values: torch.Tensor, rtol: float = 1e-8, maxiter: int = 100000,
                x0: Optional[torch.Tensor] = None):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return pcg_raw(prob.csr, values, prob.b, x0=x0, rtol=rtol, maxiter=maxiter)


def make_samples(prob: PoissonCube, seeds: Sequence[int], snapshots: Sequence[int] = (0, 8, 32, 128)) -> Dict:
    """κ fields and (residual, error) pairs from Jacobi-PCG trajectories."""
    kap, res, err, jits = [], [], [], []
    t0 = time.perf_counter()
    for i, sd in enumerate(seeds):
        kappa = prob.kappa(sd)
        values, b = prob.system(kappa)
        xs, it, rel, ok = plain_solve(prob, values)
        rs, es = [], []
        for k in snapshots:
            if k == 0:
                xk = torch.zeros_like(xs)
            else:
                xk, _, _, _ = plain_solve(prob, values, maxiter=k)
            rk = b - _spmv_raw(prob.csr, values, xk)
            rs.append(rk.to(torch.float32))
            es.append((xs - xk).to(torch.float32))
        kap.append(kappa)
        res.append(torch.stack(rs))
        err.append(torch.stack(es))
        jits.append(it)
        if (i + 1) % max(1, len(seeds) // 10) == 0:
            log(f"samples {i + 1}/{len(seeds)} ({time.perf_counter() - t0:.0f}s, Jacobi its {it})")
    return {"kappa": kap, "r": res, "e": err, "jacobi_iters": jits}


def train(prob: PoissonCube, net: GNNPreconditioner, data: Dict, steps: int, lr: float = 2e-3,
          seed: int = 0, log_every: int = 100, time_limit: float = 1e30, noise_residuals: int = 1,
          loss: str = "cosine") -> List[float]:
    """``noise_residuals`` extra residuals per step are ``r = A e`` for white
    noise ``e`` (their exact error is ``e``), so the loss covers the rough end
    of the spectrum that plain-solver residuals leave out.  ``loss``:
    ``"cosine"`` (default) is the A-norm error left after the optimal step
    along the correction, ``1 - (y.r~)^2 / ((y.A^y)(e~.r~))`` - invariant to the
    scale of M, as flexible CG is; ``"energy"`` is the relative energy error
    of the correction itself."""
    prog = net.plan(prob.hier)
    opt = Adam(net, lr=lr)
    grad = torch.zeros_like(net.params)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=prob.device)
    gen.manual_seed(seed)
    hist, window = [], []
    n = prob.n_inner
    d0 = prob.hier.levels[0]
    t0 = time.perf_counter()
    for step in range(steps):
        j = int(rng.integers(len(data["kappa"])))
        kappa = data["kappa"][j]
        values, _ = prob.system(kappa)
        prob.hier.set_matrix(values)
        lk = prob.log_kappa_feature(kappa)
        grad.zero_()
        sums = []
        pairs = [(data["r"][j][k].to(torch.float64), data["e"][j][k].to(torch.float64))
                 for k in range(data["r"][j].shape[0])]
        for _ in range(noise_residuals):
            e = torch.randn(n, generator=gen, dtype=torch.float64, device=prob.device)
            pairs.append((_spmv_raw(prob.csr, values, e), e))
        for r, e in pairs:
            # normaliser e~.r~ = e.r / s^2, s^2 = mean (d r)^2; scale g_y by 2 n / (e~.r~)
            s2 = torch.mean((d0.d.to(torch.float64) * r) ** 2)
            er = torch.dot(e, r) / s2
            if loss == "cosine":
                sm = prog.loss_and_grad(r, lk, grad, e, loss="cosine")
            else:
                sm = prog.loss_and_grad(r, lk, grad, e, gy_scale=(2.0 * n) / er)
            sums.append(sm)
        frac = step / max(1, steps - 1)
        opt.step(grad, lr=lr * 0.5 * (1.0 + math.cos(math.pi * frac)))
        S = torch.stack(sums)
        if loss == "cosine":
            window.append(1.0 - S[:, 1] ** 2 / (S[:, 0] * S[:, 2]))
        else:
            window.append((S[:, 0] - 2.0 * S[:, 1] + S[:, 2]) / S[:, 2])
        if (step + 1) % log_every == 0 or step == steps - 1:
            E = torch.cat(window).mean().item()
            hist.append(E)
            window = []
            log(f"step {step + 1}/{steps}: {loss} loss {E:.4f} ({time.perf_counter() - t0:.0f}s)")
            if time.perf_counter() - t0 > time_limit:
                log("time limit reached")
                break
    return hist


def evaluate(prob: PoissonCube, net: GNNPreconditioner, seeds: Sequence[int], chunk: int = 8,
             rtol: float = 1e-8, m: int = 4) -> Dict:
    """Time to ``rtol`` of Jacobi-PCG and of the learned FCG(m) per held-out seed."""
    solver = LearnedFCG(net, prob.hier, prob.csr, chunk=chunk, m=m)
    dev = prob.device
    # warm-up on a seed outside the evaluated range (kernel loading, graph capture)
    kappa = prob.kappa(999)
    values, b = prob.system(kappa)
    solver.set_system(values, prob.log_kappa_feature(kappa))
    solver.solve(b, rtol=rtol)
    plain_solve(prob, values, rtol=rtol, maxiter=50)
    rows = []
    for sd in seeds:
        kappa = prob.kappa(sd)
        values, b = prob.system(kappa)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        xj, itj, relj, okj = plain_solve(prob, values, rtol=rtol)
        torch.cuda.synchronize(dev)
        tj = time.perf_counter() - t0
        t0 = time.perf_counter()
        solver.set_system(values, prob.log_kappa_feature(kappa))
        xl, info = solver.solve(b, rtol=rtol)
        tl = time.perf_counter() - t0
        rows.append({"seed": int(sd), "jacobi_iters": itj, "jacobi_s": tj, "fcg_iters": info["fcg_iterations"],
                     "fcg_exit": info["fcg_exit"], "fallback_iters": info["fallback_iterations"],
                     "fcg_s": tl, "fcg_relres": info["relres"], "speedup": tj / tl})
    sp = np.array([r["speedup"] for r in rows])
    return {"rows": rows, "median_speedup": float(np.median(sp)) if len(sp) else None,
            "p10_speedup": float(np.percentile(sp, 10)) if len(sp) else None,
            "min_speedup": float(sp.min()) if len(sp) else None,
            "median_jacobi_iters": float(np.median([r["jacobi_iters"] for r in rows])) if rows else None,
            "median_fcg_iters": float(np.median([r["fcg_iters"] for r in rows])) if rows else None,
            "fallbacks": int(sum(r["fallback_iters"] > 0 for r in rows))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--train", type=int, default=2000, help="training seeds (from 1000)")
    ap.add_argument("--test", type=int, default=200, help="held-out seeds (from 5000)")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--schedule", type=str, default="1,2,3,3,2,1")
    ap.add_argument("--nonsymmetric", action="store_true",
                    help="independent up-path layers and decoder (the operator is then not SPD)")
    ap.add_argument("--n-linear", type=int, default=64)
    ap.add_argument("--bottom", type=int, default=-1, help="deepest pooled level (default: deepest layer level)")
    ap.add_argument("--general", action="store_true",
                    help="train biases and the log-coefficient input too (not homogeneous in r)")
    ap.add_argument("--noise-residuals", type=int, default=1)
    ap.add_argument("--loss", choices=["cosine", "energy"], default="cosine")
    ap.add_argument("--levels", type=int, default=0, help="hierarchy levels (default: what the schedule needs)")
    ap.add_argument("--chunk", type=int, default=8)
    ap.add_argument("--fcg-m", type=str, default="4", help="directions kept by flexible CG (comma list to compare)")
    ap.add_argument("--init", type=str, default="")
    ap.add_argument("--out", type=str, default="")
    ap.add_argument("--report", type=str, default="")
    ap.add_argument("--time-limit", type=float, default=1e30, help="seconds of training before stopping")
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    schedule = tuple(int(s) for s in args.schedule.split(","))
    t0 = time.perf_counter()
    net = (GNNPreconditioner.load(args.init, dev) if args.init
           else GNNPreconditioner(schedule, args.n_linear, dev, homogeneous=not args.general,
                                  symmetric=not (args.nonsymmetric or args.general),
                                  bottom=None if args.bottom < 0 else args.bottom))
    prob = PoissonCube(args.n, dev, n_levels=max(args.levels, net.n_levels))
    log(f"{args.n}^3: {prob.n_inner} unknowns, levels {[l.n for l in prob.hier.levels]} "
        f"({time.perf_counter() - t0:.1f}s)")
    report = {"n": args.n, "schedule": list(net.schedule), "n_linear": net.n_linear,
              "homogeneous": net.homogeneous, "symmetric": net.symmetric, "bottom": net.bottom}
    if args.train > 0 and args.steps > 0:
        data = make_samples(prob, list(range(1000, 1000 + args.train)))
        report["train_median_jacobi_iters"] = float(np.median(data["jacobi_iters"]))
        report["loss"] = args.loss
        report["loss_history"] = train(prob, net, data, args.steps, lr=args.lr, time_limit=args.time_limit,
                                       noise_residuals=args.noise_residuals, loss=args.loss)
        del data
        if args.out:
            net.save(args.out)
            log(f"saved {args.out}")
    if args.test > 0:
        for m in [int(v) for v in args.fcg_m.split(",")]:
            ev = evaluate(prob, net, list(range(5000, 5000 + args.test)), chunk=args.chunk, m=m)
            report[f"eval_m{m}"] = ev
            log(f"held-out FCG({m}): median speedup {ev['median_speedup']:.3f}, Jacobi its "
                f"{ev['median_jacobi_iters']}, FCG its {ev['median_fcg_iters']}, fallbacks {ev['fallbacks']}")
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    return report


if __name__ == "__main__":
    main()
