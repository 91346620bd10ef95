"""Batched ELN sweep: many independent runs of uneven size in one batch.

The 256-run workload: nz drawn log-uniformly from [4096, 32768] and rounded
to a multiple of 64, nvz in {64, 128} with equal probability,
sigma_nu, sigma_nubar ~ U(0.4, 0.8), alpha ~ U(0.7, 1.1), 1e-6 random-phase
seeds with per-run SplitMix64 seeds; z in [-600, 600], CFL 0.4, mu = 1,
omega = 0, FV (`--scheme fd`: the same runs with the FD scheme).  (The z
range and step count are not fixed by the workload's definition; these are
the recorded choices.)

    python -m cosenu_amd.sweep [--runs 256] [--steps 100] [--check 8] [--autotune] [--scheme fv|fd]

prints one JSON line: aggregate cell-updates/s of the batch and the
agreement of `--check` randomly chosen runs with the same runs advanced
alone (bit-identical expected).  `--autotune` first times the batch's
work-item chunk lengths on a scratch batch (cosenu_batch_autotune) and the
batch uses the fastest.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import COLL, FAST, FD, FV, STATE, STATE0, VAC, Batch, RunParams, Simulation, batch_autotune
from .initial import seeded_ic


def sweep_runs(n: int = 256, seed: int = 2026, scheme: int = FV):
    rng = np.random.default_rng(seed)
    runs, ics = [], []
    for i in range(n):
        nz = int(round(math.exp(rng.uniform(math.log(4096), math.log(32768))) / 64)) * 64
        nvz = int(rng.choice([64, 128]))
        p = RunParams(nz=nz, nvz=nvz, z0=-600.0, z1=600.0, CFL=0.4, scheme=scheme, flags=VAC | COLL,
                      omega=0.0, theta=1e-6, pmo=1.0, mu=1.0, mode=FAST)
        ics.append((float(rng.uniform(0.4, 0.8)), float(rng.uniform(0.4, 0.8)),
                    float(rng.uniform(0.7, 1.1)), int(seed * 7919 + i)))
        runs.append(p)
    return runs, ics


def run_sweep(n_runs: int = 256, steps: int = 100, warmup: int = 16, check: int = 8,
              scheme: str = "fv", autotune: bool = False) -> dict:
    """Advance the sweep's runs in one batch: `warmup` untimed steps, then
    `steps` timed ones; then `check` randomly chosen runs advanced alone for
    the same number of steps and compared with the batch bit for bit."""
    runs, ics = sweep_runs(n_runs, scheme=FD if scheme == "fd" else FV)
    tuned = None
    if autotune:
        t0 = time.perf_counter()
        L, ms = batch_autotune(runs)
        tuned = {"z_chunk": L, "ms_per_step": ms, "seconds": time.perf_counter() - t0}
    t0 = time.perf_counter()
    B = Batch(runs)
    for r, (p, (sn, sb, al, sd)) in enumerate(zip(runs, ics)):
        G, bG, s0 = seeded_ic(p, sn, sb, al, sd)
        B.set_profile(r, G[:, :1], bG[:, :1])
        B.upload(r, s0, STATE0)
        B.upload(r, s0, STATE)
    t_setup = time.perf_counter() - t0
    info = B.info()
    B.step(warmup)
    B.sync()
    t0 = time.perf_counter()
    B.step(steps)
    B.sync()
    el = time.perf_counter() - t0
    value = info["cells"] * steps / el

    # agreement with the same runs advanced alone
    rng = np.random.default_rng(1)
    pick = sorted(rng.choice(len(runs), size=min(check, len(runs)), replace=False).tolist())
    nsteps = warmup + steps
    same = 0
    for r in pick:
        p = runs[r]
        sn, sb, al, sd = ics[r]
        G, bG, s0 = seeded_ic(p, sn, sb, al, sd)
        with Simulation(p) as sim:
            sim.upload(s0)
            sim.step(nsteps)
            solo = sim.download()
        same += int(np.array_equal(solo, B.download(r)))
    B.close()
    return {"metric": "batched sweep cell-updates/s", "value": value, "unit": "cell-updates/s",
            "runs": len(runs), "scheme": scheme, "fast_variant": info["fast_variant"],
            "cells": info["cells"], "work_items": info["work_items"],
            "groups": info["ngroups"], "z_chunk": info["z_chunk"], "steps": steps,
            "warmup": warmup, "ms_per_step": 1e3 * el / steps, "setup_s": t_setup,
            "device_GB": info["device_bytes"] / 1e9,
            "checked_runs": pick, "bit_identical_to_solo": f"{same}/{len(pick)}",
            "all_checked_identical": same == len(pick), "autotune": tuned}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=256)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--check", type=int, default=8)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--scheme", choices=["fv", "fd"], default="fv")
    args = ap.parse_args(argv)
    out = run_sweep(args.runs, args.steps, args.warmup, args.check, args.scheme, args.autotune)
    print(json.dumps(out), flush=True)
    return 0 if out["all_checked_identical"] else 1


if __name__ == "__main__":
    sys.exit(main())
