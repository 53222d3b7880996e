"""Train the learned Schur smoother on seeded RT1-P0 Darcy systems, on the device.

    python -m nfempy_amd.train_learned --n 32 --systems 2000 --epochs 2 \
        --out nfempy_amd/data/learned_schur_rt1p0.npz

Data: for each training seed s (0 .. systems-1) the permeability is
``K = exp(g)``, ``g`` a Gaussian random field (squared-exponential covariance,
correlation length ``--ell``, standard deviation ``--sigma``) sampled at the cell
centroids with seed s. The system is assembled by the library's kernels and solved
by the plain method (block-diagonal preconditioner, MINRES, FP64 refinement); the
pressure errors ``p* - p_k`` of the plain Krylov iterates after ``--iterates``
iterations are kept as training samples next to Gaussian vectors.

Training: Adam on the FP32 master weights, one system per step, the two-grid
loss and its gradient from the library's loss kernels and network backward
kernels (``SmootherTrainer``). Held-out systems (seeds 10000 ..) are evaluated after
every epoch: two-grid loss and FGMRES iterations to 1e-8, with the network and
with the Jacobi smoother it replaces.
"""

from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--systems", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--ell", type=float, default=0.05)
    ap.add_argument("--sigma", type=float, default=2.0)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-min", type=float, default=1e-4)
    ap.add_argument("--random-samples", type=int, default=12)
    ap.add_argument("--iterates", type=str, default="30,300")
    ap.add_argument("--val-seeds", type=str, default="10000,10001,10002")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", type=str, default=None)
    args = ap.parse_args(argv)

    import torch

    from .batch import darcy_space
    from .device import CoefficientArg, DarcySystem
    from .fields import grf_device
    from .learned import LearnedSchurPC, SmootherTrainer, default_weights_path, init_params, solve_learned
    from .solvers import MinresIR, darcy_preconditioner

    t_start = time.time()

    def log(msg):
        print(f"[train {time.time() - t_start:7.1f}s] {msg}", flush=True)

    sysd = DarcySystem(darcy_space(args.n, 1))
    cent = sysd.coords64[sysd.cells.long()].mean(dim=1)
    f = CoefficientArg(torch.ones(1, dtype=torch.float64, device="cuda"), 0, 0)
    n_u = sysd.n_u
    E = sysd.n - n_u
    log(f"mesh {args.n}^3x6: E={E} N={sysd.n}")

    def assemble(logk):
        kinv = CoefficientArg(torch.exp(-logk).contiguous(), 1, 0)
        sysd.assemble_elements(kinv, f, fp64=True)
        sysd.assemble_elements(kinv, f, fp64=False)

    # ---- data: systems solved by the plain method ----------------------------------
    iterates = [int(v) for v in args.iterates.split(",") if v]
    logks, errs = [], []
    t_plain, it_plain = 0.0, 0
    for s in range(args.systems):
        logk = grf_device(cent, seed=s, ell=args.ell, sigma=args.sigma).to(torch.float32)
        assemble(logk.double())
        b = -sysd.rhs64
        sign, pinv = darcy_preconditioner(sysd.rowptr, sysd.col, sysd.vals64, sysd.n, n_u)
        solver = MinresIR(sysd.rowptr, sysd.col, sysd.vals32, sysd.vals64, sign, pinv)
        t0 = time.perf_counter()
        xs, info = solver.solve(b, rtol=1e-8, max_inner=100000)
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t0
        it_plain += info["iterations"]
        e = []
        for k in iterates:
            xk, _ = solver.solve(b, rtol=1e-30, max_inner=k, max_outer=1)
            e.append((xs[n_u:] - xk[n_u:]).to(torch.float32))
        solver.close()
        logks.append(logk)
        errs.append(torch.stack(e) if e else None)
        if s % 100 == 0 or s == args.systems - 1:
            log(f"data {s + 1}/{args.systems}: plain solve {info['iterations']} it, residual {info['residual']:.1e}")
    log(f"plain solves: mean {t_plain / args.systems * 1e3:.1f} ms, {it_plain / args.systems:.0f} iterations")

    # ---- training -------------------------------------------------------------------
    pc = LearnedSchurPC(sysd, None)
    tr = SmootherTrainer(pc, init_params(args.seed), lr=args.lr, n_random=args.random_samples, seed=args.seed)
    total = args.epochs * args.systems
    rng = np.random.default_rng(args.seed)
    val_seeds = [int(v) for v in args.val_seeds.split(",") if v]
    val_logk = [grf_device(cent, seed=s, ell=args.ell, sigma=args.sigma) for s in val_seeds]
    history = []

    def validate(tag):
        rows = []
        theta = pc.theta32.clone()
        for s, lk in zip(val_seeds, val_logk):
            assemble(lk)
            pc.set_params(theta)
            l_net = tr.evaluate(sysd.vals64, network=True)
            _, i_net = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8, restart=80, reorth=False)
            l_jac = tr.evaluate(sysd.vals64, network=False)
            _, i_jac = solve_learned(sysd, pc, -sysd.rhs64, rtol=1e-8, restart=80, reorth=False)
            rows.append({"seed": s, "loss_net": l_net, "loss_jacobi": l_jac, "it_net": i_net["iterations"],
                         "it_jacobi": i_jac["iterations"], "fallback": i_net["fallback"]})
        pc.set_params(theta)
        tr.opt.theta = pc.theta32
        mean = {k: float(np.mean([r[k] for r in rows])) for k in ("loss_net", "loss_jacobi", "it_net", "it_jacobi")}
        log(f"validate {tag}: loss net {mean['loss_net']:.4f} jacobi {mean['loss_jacobi']:.4f}; "
            f"FGMRES it net {mean['it_net']:.1f} jacobi {mean['it_jacobi']:.1f}")
        return {"tag": tag, "rows": rows, **mean}

    history.append(validate("init"))
    step = 0
    for ep in range(args.epochs):
        order = rng.permutation(args.systems)
        run = []
        for s in order:
            lr = args.lr_min + 0.5 * (args.lr - args.lr_min) * (1 + math.cos(math.pi * step / max(total - 1, 1)))
            tr.opt.lr = lr
            assemble(logks[s].double())
            loss = tr.step(sysd.vals64, extra=errs[s])
            run.append(loss)
            step += 1
            if step % 200 == 0:
                log(f"epoch {ep} step {step}/{total} lr {lr:.2e} loss {np.mean(run[-200:]):.4f}")
        history.append(validate(f"epoch{ep}"))

    theta = pc.theta32.detach().cpu().numpy().astype(np.float32)
    out = args.out or str(default_weights_path())
    meta = {"n": args.n, "systems": args.systems, "train_seeds": [0, args.systems - 1], "epochs": args.epochs,
            "ell": args.ell, "sigma": args.sigma, "iterates": iterates, "random_samples": args.random_samples,
            "lr": [args.lr, args.lr_min], "plain_ms": t_plain / args.systems * 1e3,
            "plain_iterations": it_plain / args.systems, "history": history}
    np.savez(out, theta=theta, meta=np.array(json.dumps(meta)))
    log(f"saved {out}")
    print(json.dumps({"final": history[-1]}), flush=True)


if __name__ == "__main__":
    main()
