"""A long single run with asynchronous P_ee snapshots and conservation checks.

    python -m cosenu_amd.large_run --workload c4 --steps 5000 --snap-every 500

Advances a named workload in FAST mode and, every --snap-every steps, copies
the P_ee(z, v) = ee/ee0 plane to page-locked host memory on the copy stream
(cosenu_snapshot_pee_async: the copy of one snapshot overlaps the steps that
follow it) and evaluates the device-side diagnostics: total trace, lepton
number and |M0| with their drift from t = 0, <P_ee>, <P_ee-bar>, and the
non-finite count.  The host summarises each snapshot (mean, min, max of
P_ee, finiteness) while the device keeps stepping.  Prints one JSON line per
snapshot and a final line with the throughput including snapshots.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import FAST, FD, FV, STATE, STATE0, PinnedBuffer, Simulation
from .initial import workload


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--scheme", choices=["fv", "fd"], default="fv")
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--snap-every", type=int, default=500)
    args = ap.parse_args(argv)

    wl = workload(args.workload, mode=FAST, scheme=FV if args.scheme == "fv" else FD)
    p = wl.params
    t_setup = time.perf_counter()
    a, b, G1, bG1 = wl.factors()
    sim = Simulation(p)
    sim.set_profile_v(G1, bG1)
    sim.init_separable(a, b, STATE0)
    sim.init_separable(a, b, STATE)
    d0 = sim.diagnostics()
    cells = p.nz * p.nvz
    bufs = [PinnedBuffer(cells), PinnedBuffer(cells)]
    print(json.dumps({"workload": wl.name, "nz": p.nz, "nvz": p.nvz, "cells": cells,
                      "device_bytes": sim.info()["device_bytes"],
                      "setup_s": time.perf_counter() - t_setup, "t0": d0}), flush=True)

    def summarise(buf, step):
        x = buf.array
        return {"snapshot_step": step, "pee_mean": float(x.mean()), "pee_min": float(x.min()),
                "pee_max": float(x.max()), "finite": bool(np.isfinite(x).all())}

    pending = None  # (buffer index, step) of the snapshot in flight
    done = 0
    k = 0
    records = []
    t0 = time.perf_counter()
    while done < args.steps:
        n = min(args.snap_every, args.steps - done)
        sim.step(n)
        done += n
        if pending is not None:
            sim.snapshot_wait()  # the copy of the previous snapshot has landed
        d = sim.diagnostics()   # synchronises the compute stream
        sim.snapshot_pee_async(bufs[k % 2])
        rec = {"step": done, "elapsed_s": time.perf_counter() - t0,
               "trace_drift": abs(d["trace"] - d0["trace"]) / abs(d0["trace"]),
               "btrace_drift": abs(d["btrace"] - d0["btrace"]) / abs(d0["btrace"]),
               "lepton_drift": abs(d["lepton"] - d0["lepton"]) / abs(d0["lepton"]),
               "M0_norm": d["M0_norm"], "Pee": d["Pee"], "Pbee": d["Pbee"],
               "dP_max": d["dP_max"], "nonfinite": d["nonfinite"]}
        if pending is not None:  # summarise the previous snapshot while this one copies
            rec["previous_snapshot"] = summarise(bufs[pending[0]], pending[1])
        pending = (k % 2, done)
        k += 1
        records.append(rec)
        print(json.dumps(rec), flush=True)
    sim.snapshot_wait()
    el = time.perf_counter() - t0
    last = summarise(bufs[pending[0]], pending[1])
    out = {"workload": wl.name, "scheme": args.scheme, "steps": args.steps,
           "snap_every": args.snap_every, "elapsed_s": el,
           "cell_updates_per_s_with_snapshots": cells * args.steps / el,
           "last_snapshot": last, "final": records[-1],
           "max_trace_drift": max(r["trace_drift"] for r in records),
           "max_lepton_drift": max(r["lepton_drift"] for r in records)}
    print(json.dumps(out), flush=True)
    for bf in bufs:
        bf.free()
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
