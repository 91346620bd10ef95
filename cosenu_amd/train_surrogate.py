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

The dataset can be produced in pieces (one GPU job each):

    python -m cosenu_amd.train_surrogate --nz 8192 --steps 8000 --runs 2048 \
        --data-dir D --gen 0 1024
    ... --gen 1024 2048
    python -m cosenu_amd.train_surrogate --nz 8192 --steps 8000 --runs 2048 --data-dir D

Run i of the sample is the same whatever the sample size, so a piece
`data_<a>_<b>.npz` holds runs [a, b) of every sample with the same seed; the
full job loads the pieces it finds (same grid, schedule and seed) and solves
what is missing.
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


def _meta(args, setup):
    return {"nz": setup.nz, "z0": setup.z0, "z1": setup.z1, "steps": setup.steps,
            "avg_frac": setup.avg_frac, "sample_every": setup.sample_every, "nvz": S.NV,
            "seed": args.seed}


def save_piece(args, setup, a, b):
    """Solve runs [a, b) of the sample and write them to --data-dir."""
    elns = S.sample_elns(b, args.seed)[a:b]
    X = np.stack([S.features(e) for e in elns])
    Y, t = S.solve(elns, setup, log=log)
    os.makedirs(args.data_dir, exist_ok=True)
    path = os.path.join(args.data_dir, f"data_{a}_{b}.npz")
    np.savez_compressed(path, X=X, Y=Y, a=a, b=b, solver_s=t,
                        meta=json.dumps(_meta(args, setup)))
    log(json.dumps({"piece": [a, b], "solver_s": t, "file": path}))


def dataset(args, setup, elns):
    """Targets of runs [0, len(elns)): pieces found in --data-dir with the same
    grid, schedule and seed are loaded, the rest solved.
    Returns (Y, solver seconds, [[a, b, source], ...])."""
    n = len(elns)
    Y = np.full((n, S.NV), np.nan, dtype=np.float32)
    have = np.zeros(n, dtype=bool)
    t = 0.0
    pieces = []
    meta = _meta(args, setup)
    if args.data_dir and os.path.isdir(args.data_dir):
        for name in sorted(os.listdir(args.data_dir)):
            if not (name.startswith("data_") and name.endswith(".npz")):
                continue
            with np.load(os.path.join(args.data_dir, name)) as z:
                if json.loads(str(z["meta"])) != meta:
                    continue
                a, b = int(z["a"]), int(z["b"])
                if a >= n:
                    continue
                m = min(b, n) - a
                X = np.stack([S.features(e) for e in elns[a:a + m]])
                if not np.array_equal(X, z["X"][:m]):
                    raise SystemExit(f"{name}: features differ from the sample")
                Y[a:a + m] = z["Y"][:m]
                have[a:a + m] = True
                t += float(z["solver_s"]) * m / (b - a)
                pieces.append([a, a + m, name])
    todo = np.nonzero(~have)[0]
    if todo.size:
        Yt, tt = S.solve([elns[i] for i in todo], setup, log=log)
        Y[todo] = Yt
        t += tt
        cut = np.nonzero(np.diff(todo) > 1)[0] + 1
        for seg in np.split(todo, cut):
            pieces.append([int(seg[0]), int(seg[-1]) + 1, "solved here"])
    return Y, t, pieces


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
    ap.add_argument("--data-dir", default="", help="directory of dataset pieces data_<a>_<b>.npz")
    ap.add_argument("--gen", type=int, nargs=2, metavar=("A", "B"),
                    help="solve runs [A, B) of the sample into --data-dir and exit")
    args = ap.parse_args(argv)

    setup = S.SolverSetup(nz=args.nz, z0=-args.zhalf, z1=args.zhalf, steps=args.steps)
    t0 = time.perf_counter()
    if args.gen:
        a, b = args.gen
        if not args.data_dir or not 0 <= a < b:
            ap.error("--gen A B needs --data-dir and 0 <= A < B")
        save_piece(args, setup, a, b)
        return 0
    elns = S.sample_elns(args.runs, args.seed)
    X = np.stack([S.features(e) for e in elns])
    Y, t_data, pieces = dataset(args, setup, elns)
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
                    "avg_over_final_fraction": setup.avg_frac,
                    "sample_every_steps": setup.sample_every, "solver_s": t_data,
                    "pieces": pieces},
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
