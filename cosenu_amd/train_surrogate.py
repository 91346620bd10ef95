"""End-to-end surrogate pipeline on one GPU.

    python -m cosenu_amd.train_surrogate [--runs 2000] [--sweep 256] [--out DIR]

1. Dataset: `--runs` ELN configurations (sigma_nu, sigma_nubar ~ U(0.4, 0.8),
   alpha ~ U(0.7, 1.1)) advanced by the batched FP64 solver; targets are the
   time-averaged P_ee(v) over the final 20% of the steps.  90/10 split.
2. Training: an ensemble of networks (67 -> 256 -> 256 -> 256 -> 64) trained
   with the library's BF16-MFMA forward/backward/AdamW kernels.
3. Calibration: the acceptance threshold on the ensemble spread that keeps the
   held-out mean |dP_ee| within 2%.
4. Screened sweep: `--sweep` fresh ELN configurations.  Solver-only: the
   batched solver runs all of them.  Screened: the ensemble predicts all of
   them and the solver runs only the rejected ones.  Reports the speed-up and
   the mean |dP_ee| of the screened result against the solver's.
5. Ablation: the trained ensemble against the method without a network
   (predicting the training-set mean P_ee(v)).

Prints one JSON line and writes it (plus the ensemble weights) to --out.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import surrogate as S


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=2000)
    ap.add_argument("--sweep", type=int, default=256)
    ap.add_argument("--nz", type=int, default=2048)
    ap.add_argument("--zhalf", type=float, default=600.0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--bound", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)

    setup = S.SolverSetup(nz=args.nz, z0=-args.zhalf, z1=args.zhalf, steps=args.steps)
    t0 = time.perf_counter()
    elns = S.sample_elns(args.runs, args.seed)
    X = np.stack([S.features(e) for e in elns])
    Y, t_data = S.solve(elns, setup, log=log)
    log(f"dataset: {args.runs} runs, solver {t_data:.1f} s")
    n_train = int(round(0.9 * args.runs))
    Xtr, Ytr, Xte, Yte = X[:n_train], Y[:n_train], X[n_train:], Y[n_train:]

    cfg = S.TrainConfig(steps=args.train_steps, members=args.members)
    t1 = time.perf_counter()
    members = []
    curves = []
    for k in range(cfg.members):
        net, losses = S.train_member(Xtr, Ytr, cfg, seed=1000 + k, log=log)
        members.append(net)
        curves.append(losses)
    S_ = S.Surrogate(members, bound=args.bound)
    t_train = time.perf_counter() - t1
    cal = S_.calibrate(Xte, Yte)
    log(f"calibration: {cal}")

    # method without the network: predict the training-set mean profile
    base = Ytr.mean(axis=0)
    base_err = float(np.abs(base[None, :] - Yte).mean())

    # screened sweep on fresh configurations
    fresh = S.sample_elns(args.sweep, args.seed + 1)
    Xs = np.stack([S.features(e) for e in fresh])
    Ys, t_solver_all = S.solve(fresh, setup)
    ta = time.perf_counter()
    pred, accept = S_.screen(Xs)
    t_infer = time.perf_counter() - ta
    rejected = [fresh[i] for i in np.nonzero(~accept)[0]]
    if rejected:
        Yr, t_fallback = S.solve(rejected, setup)
        pred[~accept] = Yr
    else:
        t_fallback = 0.0
    err_sweep = np.abs(pred - Ys).mean(axis=1)
    t_screened = t_infer + t_fallback
    out = {
        "metric": "surrogate-screened ELN sweep vs batched solver",
        "sweep_runs": args.sweep,
        "solver_only_s": t_solver_all,
        "screened_s": t_screened,
        "surrogate_inference_s": t_infer,
        "fallback_runs": len(rejected),
        "fallback_s": t_fallback,
        "speedup": t_solver_all / t_screened if t_screened > 0 else None,
        "sweep_mean_abs_dPee": float(err_sweep.mean()),
        "sweep_max_run_mean_abs_dPee": float(err_sweep.max()),
        "bound": args.bound,
        "heldout": cal,
        "no_network_baseline_heldout_mean_abs_dPee": base_err,
        "dataset": {"runs": args.runs, "train": n_train, "heldout": args.runs - n_train,
                    "nz": setup.nz, "nvz": S.NV, "z": [setup.z0, setup.z1], "steps": setup.steps,
                    "avg_over_final_fraction": setup.avg_frac, "solver_s": t_data},
        "training": {"members": cfg.members, "steps": cfg.steps, "batch": cfg.batch, "lr": cfg.lr,
                     "weight_decay": cfg.weight_decay, "seconds": t_train,
                     "loss_curves": curves},
        "wall_s": time.perf_counter() - t0,
    }
    line = json.dumps(out)
    print(line, flush=True)
    if args.out:
        os.makedirs(args.out, exist_ok=True)
        with open(os.path.join(args.out, "surrogate_report.json"), "w") as f:
            f.write(line + "\n")
        np.savez_compressed(os.path.join(args.out, "surrogate_weights.npz"),
                            *[m.get_params() for m in members], tau=S_.tau)
    return 0


if __name__ == "__main__":
    sys.exit(main())
