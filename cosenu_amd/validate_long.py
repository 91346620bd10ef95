"""Long-run agreement of FAST mode with STRICT mode on a full workload.

STRICT mode evaluates the reference's arithmetic bit for bit (checked against
the reference build in the tests), so on the full-size workloads, where the
CPU reference would take days, it stands in for the reference:

    python -m cosenu_amd.validate_long --workload c1 --steps 20000 [--scheme fd]

advances the same initial state in both modes and reports, every --every
steps and at the end: P_ee(z, v) = ee/ee0 normwise (||d||_inf/||P||_inf and
L2), the total trace and the lepton number of each mode, and their drift
from t = 0.  Prints one JSON line at the end.

Where the run turns unstable (the flavour instability grows), any
difference of rounding grows with it, and two bit-exact evaluations of the
reference's arithmetic drift apart as well.  --against strict-ulp measures
that intrinsic spread: it advances, in place of FAST, a second STRICT run
whose initial state differs from the first by one unit in the last place
in every value (sign random, seed fixed).  FAST's distance from STRICT is
then judged against STRICT's distance from itself.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import FAST, FD, FV, STATE, STATE0, STRICT, Simulation
from .initial import workload


def subsampled(wl, nz: int):
    """The workload on a domain of nz cells with the same dz, centred on z = 0
    (boundary, physics, initial-state recipe and seed unchanged)."""
    p = wl.params
    half = 0.5 * nz * (p.z1 - p.z0) / p.nz
    wl.params = dataclasses.replace(p, nz=nz, z0=-half, z1=half)
    return wl


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c1")
    ap.add_argument("--scheme", choices=["fv", "fd"], default="fv")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--every", type=int, default=2000)
    ap.add_argument("--nz", type=int, default=0,
                    help="z-subsampled domain: this many cells at the workload's dz")
    ap.add_argument("--against", choices=["fast", "strict-ulp"], default="fast",
                    help="the second run: FAST mode (default), or STRICT mode from the "
                         "initial state perturbed by one ulp per value")
    args = ap.parse_args(argv)
    scheme = FV if args.scheme == "fv" else FD

    sims = {}
    second = FAST if args.against == "fast" else STRICT
    for name, mode in (("strict", STRICT), ("fast", second)):
        wl = workload(args.workload, mode=mode, scheme=scheme)
        if args.nz:
            wl = subsampled(wl, args.nz)
        a, b, G1, bG1 = wl.factors()
        sim = Simulation(wl.params)
        sim.set_profile_v(G1, bG1)
        sim.init_separable(a, b, STATE0)
        sim.init_separable(a, b, STATE)
        if name == "fast" and args.against == "strict-ulp":
            y = sim.download()
            sgn = np.where(np.random.default_rng(12345).random(y.shape) < 0.5, -1.0, 1.0)
            sim.upload(np.where(y != 0.0, y + sgn * np.spacing(y), y))
        sims[name] = sim
    p = wl.params
    ee0 = sims["strict"].download(STATE0)[0]
    d0 = {k: s.diagnostics() for k, s in sims.items()}
    history = []
    done = 0
    t0 = time.perf_counter()
    while done < args.steps:
        n = min(args.every, args.steps - done)
        for s in sims.values():
            s.step(n)
        for s in sims.values():
            s.sync()
        done += n
        st = sims["strict"].download()
        fa = sims["fast"].download()
        ps, pf = st[0] / ee0, fa[0] / ee0
        linf = float(np.max(np.abs(pf - ps)) / np.max(np.abs(ps)))
        l2 = float(np.linalg.norm(pf - ps) / np.linalg.norm(ps))
        state_linf = float(np.max(np.abs(fa - st)) / np.max(np.abs(st)))
        d = {k: s.diagnostics() for k, s in sims.items()}
        rec = {"step": done, "pee_linf": linf, "pee_l2": l2, "state_linf": state_linf,
               "trace_rel_diff": abs(d["fast"]["trace"] - d["strict"]["trace"]) / abs(d["strict"]["trace"]),
               "lepton_rel_diff": abs(d["fast"]["lepton"] - d["strict"]["lepton"]) / abs(d["strict"]["lepton"]),
               "trace_drift_strict": abs(d["strict"]["trace"] - d0["strict"]["trace"]) / abs(d0["strict"]["trace"]),
               "trace_drift_fast": abs(d["fast"]["trace"] - d0["fast"]["trace"]) / abs(d0["fast"]["trace"]),
               "lepton_drift_strict": abs(d["strict"]["lepton"] - d0["strict"]["lepton"]) / abs(d0["strict"]["lepton"]),
               "lepton_drift_fast": abs(d["fast"]["lepton"] - d0["fast"]["lepton"]) / abs(d0["fast"]["lepton"]),
               "Pee_strict": d["strict"]["Pee"], "Pee_fast": d["fast"]["Pee"],
               "dP_max_strict": d["strict"]["dP_max"], "nonfinite": d["fast"]["nonfinite"] + d["strict"]["nonfinite"],
               "elapsed_s": time.perf_counter() - t0}
        history.append(rec)
        print(json.dumps({"progress": rec}), flush=True)
    last = history[-1]
    fast = sims["fast"]
    out = {"workload": args.workload, "scheme": args.scheme, "nz": p.nz, "nvz": p.nvz,
           "z": [p.z0, p.z1],
           "against": args.against,
           "fast_variant": fast.info()["fast_variant"], "fast_kernel": fast.stage_kernel_name(0),
           "range_teams": fast.info()["range_teams"],
           "steps": args.steps, "pass_1e-10": bool(last["pee_linf"] <= 1e-10 and last["pee_l2"] <= 1e-10
                                                   and last["trace_rel_diff"] <= 1e-10
                                                   and last["lepton_rel_diff"] <= 1e-10),
           "final": last, "history": history}
    print(json.dumps(out), flush=True)
    for s in sims.values():
        s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
