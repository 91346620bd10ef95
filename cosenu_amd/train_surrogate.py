"""End-to-end surrogate pipeline on one GPU.

    python -m cosenu_amd.train_surrogate --nz 8192 --steps 8000 --runs 10000 \
        --dataset cosenu_amd/data/eln_dataset_nz8192.npz --out DIR

1. Dataset: `--runs` ELN configurations (sigma_nu, sigma_nubar ~ U(0.4, 0.8),
   alpha ~ U(0.7, 1.1), per-run phase seeds) advanced by the batched FP64
   solver; targets are P_ee(v) averaged over z and over the final 20% of the
   steps.  Split 90/10: the first 90% train the networks; the last 10% are
   held out, half to calibrate the gate and half to measure it.
2. Training: an ensemble of networks (67 -> 256 -> 256 -> 256 -> 64) trained
   with the library's BF16-MFMA forward/backward/AdamW kernels.
3. Gates on the ensemble spread, calibrated on the calibration half and
   measured on the test half (accept rate, largest accepted error, recall of
   the runs above the bound): the per-run gate, below which no calibration
   run exceeds the bound (mean over bins of |dP_ee| <= 2%), and the
   sweep-mean gate, at which the screened sweep's mean error stays within
   the bound.  `--replicas` runs (the same configurations under other phase
   seeds) measure how far the solver's own targets move with the seed.
4. Screened sweep: `--sweep` fresh ELN configurations.  Solver-only: the
   batched solver runs all of them.  Screened, for each gate: the ensemble
   predicts all of them and the solver runs only the rejected ones.  Reports
   the speed-up, the mean |dP_ee| of the screened result against the
   solver's and the largest error of an accepted run.
5. Ablation: the trained ensemble against the method without a network
   (predicting the training-set mean P_ee(v)).

Prints one JSON line and writes it to DIR/surrogate_report.json, the ensemble
(weights, threshold) to DIR/surrogate_ensemble.npz.

The dataset is produced in pieces (one GPU job each), one file per solver
batch:

    python -m cosenu_amd.train_surrogate --nz 8192 --steps 8000 --runs 10000 \
        --data-dir D --gen 0 1536
    ... --gen 1536 3072

Run i of the sample is the same whatever the sample size, so a piece
`data_<a>_<b>.npz` holds runs [a, b) of every sample with the same seed.
`--collect FILE` gathers the pieces of `--data-dir` into one dataset file
(ELN parameters and targets); `--dataset FILE` reads such a file, and runs it
lacks are loaded from `--data-dir` pieces or solved.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

from . import surrogate as S

REPLICA_SEED_XOR = 0x5DEECE66D  # phase seeds of the replica runs


def log(msg):
    print(msg, flush=True)


def _meta(args, setup):
    return {"nz": setup.nz, "z0": setup.z0, "z1": setup.z1, "steps": setup.steps,
            "avg_frac": setup.avg_frac, "sample_every": setup.sample_every, "nvz": S.NV,
            "seed": args.seed}


def save_piece(args, setup, a, b):
    """Solve runs [a, b) of the sample and write them to --data-dir."""
    os.makedirs(args.data_dir, exist_ok=True)
    sample = S.sample_elns(b, args.seed)
    # one file per solver batch, written as soon as the batch is done
    for a0 in range(a, b, setup.batch):
        a1 = min(b, a0 + setup.batch)
        elns = sample[a0:a1]
        X = np.stack([S.features(e) for e in elns])
        Y, t = S.solve(elns, setup, log=log)
        path = os.path.join(args.data_dir, f"data_{a0}_{a1}.npz")
        np.savez_compressed(path, X=X, Y=Y, a=a0, b=a1, solver_s=t,
                            meta=json.dumps(_meta(args, setup)))
        log(json.dumps({"piece": [a0, a1], "solver_s": t, "file": path}))


def replica(e):
    """The same ELN configuration with another phase seed."""
    return dataclasses.replace(e, seed=e.seed ^ REPLICA_SEED_XOR)


def save_replicas(args, setup, a, b):
    """Runs [a, b) of the sample again with other phase seeds (rep_<a>_<b>.npz):
    how far the solver's own targets move with the seed of the 1e-6
    perturbation, which is not an input of the network."""
    os.makedirs(args.data_dir, exist_ok=True)
    sample = S.sample_elns(b, args.seed)
    for a0 in range(a, b, setup.batch):
        a1 = min(b, a0 + setup.batch)
        Y, t = S.solve([replica(e) for e in sample[a0:a1]], setup, log=log)
        path = os.path.join(args.data_dir, f"rep_{a0}_{a1}.npz")
        np.savez_compressed(path, Y=Y, a=a0, b=a1, solver_s=t,
                            meta=json.dumps(_meta(args, setup)))
        log(json.dumps({"replicas": [a0, a1], "solver_s": t, "file": path}))


def load_replicas(args, setup, n):
    """Replica targets of runs [0, n) found in --data-dir or the dataset file:
    (Y_rep with NaN rows where none, mask)."""
    Yr = np.full((n, S.NV), np.nan, dtype=np.float32)
    if args.dataset and os.path.exists(args.dataset):
        with np.load(args.dataset, allow_pickle=False) as z:
            if "Y_rep" in z.files and json.loads(str(z["meta"])) == _meta(args, setup):
                m = min(n, len(z["Y_rep"]))
                Yr[:m] = z["Y_rep"][:m]
    if args.data_dir and os.path.isdir(args.data_dir):
        for name in sorted(os.listdir(args.data_dir)):
            if not (name.startswith("rep_") and name.endswith(".npz")):
                continue
            with np.load(os.path.join(args.data_dir, name), allow_pickle=False) as z:
                if json.loads(str(z["meta"])) != _meta(args, setup):
                    continue
                a, b = int(z["a"]), min(int(z["b"]), n)
                if a < b:
                    Yr[a:b] = z["Y"][:b - a]
    return Yr, ~np.isnan(Yr).any(axis=1)


def seed_reproducibility(Y, Yr, have, bound):
    """Per-run mean |dP_ee| between the solver's targets of one ELN
    configuration under two phase seeds."""
    if not have.any():
        return None
    d = S.run_errors(Y[have], Yr[have])
    return {"runs": int(have.sum()), "mean": float(d.mean()), "median": float(np.median(d)),
            "q90": float(np.quantile(d, 0.9)), "q95": float(np.quantile(d, 0.95)),
            "max": float(d.max()), "fraction_above_bound": float((d > bound).mean())}


def load_dataset_file(path, args, setup, elns):
    """Targets from a collected dataset file; checks grid, schedule, seed and
    the ELN parameters of every run it holds.  Returns (Y, have, solver s)."""
    n = len(elns)
    Y = np.full((n, S.NV), np.nan, dtype=np.float32)
    have = np.zeros(n, dtype=bool)
    with np.load(path, allow_pickle=False) as z:
        if json.loads(str(z["meta"])) != _meta(args, setup):
            raise SystemExit(f"{path}: grid, schedule or seed differ")
        m = min(n, len(z["Y"]))
        eln = np.array([[e.sig_nu, e.sig_nubar, e.alpha, e.seed] for e in elns[:m]])
        if not np.array_equal(eln, z["eln"][:m]):
            raise SystemExit(f"{path}: ELN parameters differ from the sample")
        Y[:m] = z["Y"][:m]
        have[:m] = True
        t = float(z["solver_s"]) * m / len(z["Y"])
    return Y, have, t


def dataset(args, setup, elns):
    """Targets of runs [0, len(elns)): a collected dataset file (--dataset),
    then pieces found in --data-dir with the same grid, schedule and seed,
    the rest solved.  Returns (Y, solver seconds, [[a, b, source], ...])."""
    n = len(elns)
    Y = np.full((n, S.NV), np.nan, dtype=np.float32)
    have = np.zeros(n, dtype=bool)
    t = 0.0
    pieces = []
    meta = _meta(args, setup)
    if args.dataset and os.path.exists(args.dataset):
        Y, have, t = load_dataset_file(args.dataset, args, setup, elns)
        if have.any():
            pieces.append([0, int(have.sum()), os.path.basename(args.dataset)])
    if args.data_dir and os.path.isdir(args.data_dir):
        for name in sorted(os.listdir(args.data_dir)):
            if not (name.startswith("data_") and name.endswith(".npz")):
                continue
            with np.load(os.path.join(args.data_dir, name), allow_pickle=False) as z:
                if json.loads(str(z["meta"])) != meta:
                    continue
                a, b = int(z["a"]), int(z["b"])
                if a >= n or have[a:min(b, n)].all():
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


def collect(args, setup, path):
    """Gather the --data-dir pieces of runs [0, --runs) into one file."""
    elns = S.sample_elns(args.runs, args.seed)
    args_dataset, args.dataset = args.dataset, ""
    n = len(elns)
    Y = np.full((n, S.NV), np.nan, dtype=np.float32)
    t = 0.0
    found = np.zeros(n, dtype=bool)
    for name in sorted(os.listdir(args.data_dir)):
        if not (name.startswith("data_") and name.endswith(".npz")):
            continue
        with np.load(os.path.join(args.data_dir, name), allow_pickle=False) as z:
            if json.loads(str(z["meta"])) != _meta(args, setup):
                continue
            a, b = int(z["a"]), min(int(z["b"]), n)
            if a >= n:
                continue
            X = np.stack([S.features(e) for e in elns[a:b]])
            if not np.array_equal(X, z["X"][:b - a]):
                raise SystemExit(f"{name}: features differ from the sample")
            Y[a:b] = z["Y"][:b - a]
            found[a:b] = True
            t += float(z["solver_s"]) * (b - a) / (int(z["b"]) - a)
    m = int(np.argmin(found)) if not found.all() else n  # the leading complete runs
    Yr, have_r = load_replicas(args, setup, m)
    args.dataset = args_dataset
    r = int(np.argmin(have_r)) if not have_r.all() else m  # leading runs with replicas
    eln = np.array([[e.sig_nu, e.sig_nubar, e.alpha, e.seed] for e in elns[:m]])
    np.savez_compressed(path, eln=eln, Y=Y[:m], Y_rep=Yr[:r],
                        solver_s=np.float64(t * m / max(1, found.sum())),
                        meta=json.dumps(_meta(args, setup)))
    log(json.dumps({"collected_runs": m, "replica_runs": r, "file": path}))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=10000)
    ap.add_argument("--sweep", type=int, default=256)
    ap.add_argument("--nz", type=int, default=8192)
    ap.add_argument("--zhalf", type=float, default=600.0)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--bound", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default="")
    ap.add_argument("--dataset", default="", help="collected dataset file (.npz)")
    ap.add_argument("--data-dir", default="", help="directory of dataset pieces data_<a>_<b>.npz")
    ap.add_argument("--gen", type=int, nargs=2, metavar=("A", "B"),
                    help="solve runs [A, B) of the sample into --data-dir and exit")
    ap.add_argument("--gen-replicas", type=int, nargs=2, metavar=("A", "B"),
                    help="solve runs [A, B) again with other phase seeds into --data-dir and exit")
    ap.add_argument("--collect", default="", help="gather the --data-dir pieces into this file")
    args = ap.parse_args(argv)

    setup = S.SolverSetup(nz=args.nz, z0=-args.zhalf, z1=args.zhalf, steps=args.steps)
    t0 = time.perf_counter()
    if args.gen:
        a, b = args.gen
        if not args.data_dir or not 0 <= a < b:
            ap.error("--gen A B needs --data-dir and 0 <= A < B")
        save_piece(args, setup, a, b)
        return 0
    if args.gen_replicas:
        a, b = args.gen_replicas
        if not args.data_dir or not 0 <= a < b:
            ap.error("--gen-replicas A B needs --data-dir and 0 <= A < B")
        save_replicas(args, setup, a, b)
        return 0
    if args.collect:
        if not args.data_dir:
            ap.error("--collect needs --data-dir")
        collect(args, setup, args.collect)
        return 0
    elns = S.sample_elns(args.runs, args.seed)
    X = np.stack([S.features(e) for e in elns])
    Y, t_data, pieces = dataset(args, setup, elns)
    log(f"dataset: {args.runs} runs, solver {t_data:.1f} s")
    Yr, have_r = load_replicas(args, setup, args.runs)
    seed_noise = seed_reproducibility(Y, Yr, have_r, args.bound)
    log(f"seed reproducibility: {seed_noise}")
    n_train = int(round(0.9 * args.runs))
    n_cal = (args.runs - n_train) // 2
    Xtr, Ytr = X[:n_train], Y[:n_train]
    Xca, Yca = X[n_train:n_train + n_cal], Y[n_train:n_train + n_cal]
    Xte, Yte = X[n_train + n_cal:], Y[n_train + n_cal:]

    cfg = S.TrainConfig(steps=args.train_steps, members=args.members)
    t1 = time.perf_counter()
    members = []
    curves = []
    for k in range(cfg.members):
        net, losses = S.train_member(Xtr, Ytr, cfg, seed=1000 + k, log=log)
        members.append(net)
        curves.append(losses)
    sweep_seed = args.seed + 1
    S_ = S.Surrogate(members, bound=args.bound,
                     meta={"dataset": _meta(args, setup), "train_runs": n_train,
                           "calibration_runs": [n_train, n_train + n_cal],
                           "sweep_seed": sweep_seed})
    t_train = time.perf_counter() - t1
    cal = S_.calibrate(Xca, Yca)
    log(f"calibration: {cal}")
    test = S_.evaluate(Xte, Yte)
    log(f"test: {test}")

    # method without the network: predict the training-set mean profile
    base = Ytr.mean(axis=0)
    base_err = float(S.run_errors(base[None, :], Yte).mean())

    # screened sweep on fresh configurations
    fresh = S.sample_elns(args.sweep, sweep_seed)
    Xs = np.stack([S.features(e) for e in fresh])
    log("sweep: solver on every run")
    Ys, t_solver_all = S.solve(fresh, setup, log=log)
    sweeps = {}
    for gate in ("run", "mean"):
        ta = time.perf_counter()
        pred, accept = S_.screen(Xs, gate=gate)
        t_infer = time.perf_counter() - ta
        err_net = S.run_errors(pred, Ys)
        rejected = [fresh[i] for i in np.nonzero(~accept)[0]]
        if rejected:
            log(f"sweep, {gate} gate: solver on the {len(rejected)} rejected runs")
            Yr, t_fallback = S.solve(rejected, setup)
            pred[~accept] = Yr
        else:
            t_fallback = 0.0
        err_sweep = S.run_errors(pred, Ys)
        t_screened = t_infer + t_fallback
        sweeps[gate] = {
            "screened_s": t_screened, "surrogate_inference_s": t_infer,
            "fallback_runs": len(rejected), "fallback_s": t_fallback,
            "speedup": t_solver_all / t_screened if t_screened > 0 else None,
            "sweep_mean_abs_dPee": float(err_sweep.mean()),
            "sweep_max_accepted_run_abs_dPee": float(err_net[accept].max()) if accept.any() else None,
            "accepted_runs_above_bound": int(((err_net > args.bound) & accept).sum()),
            "sweep_max_run_abs_dPee_network_only": float(err_net.max()),
            "sweep_mean_abs_dPee_network_only": float(err_net.mean())}
        log(f"sweep, {gate} gate: {sweeps[gate]}")
    out = {
        "metric": "surrogate-screened ELN sweep vs batched solver",
        "sweep_runs": args.sweep,
        "solver_only_s": t_solver_all,
        "per_run_gate": sweeps["run"],
        "sweep_mean_gate": sweeps["mean"],
        "bound": args.bound,
        "calibration": cal,
        "heldout": test,
        "no_network_baseline_heldout_mean_abs_dPee": base_err,
        "solver_seed_reproducibility": seed_noise,
        "dataset": {"runs": args.runs, "train": n_train, "calibration": n_cal,
                    "test": args.runs - n_train - n_cal,
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
        S_.save(os.path.join(args.out, "surrogate_ensemble.npz"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
