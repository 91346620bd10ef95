#!/usr/bin/env python3
"""Benchmark of the COSEnu time-step hot path on one MI355X.

Workload (the single-GPU production configuration): FV-WENO7 + RK4, FP64,
nz = 65536 cells over z in [-1500, 1500], nvz = 128 velocity bins, CFL 0.4,
periodic, mu = 1, omega = 0; synthetic flavour-unstable initial state
(sigma_nu = 0.6, sigma_nubar = 0.53, alpha = 0.92, 1e-6 random-phase seed plus
a Gaussian seed at z = 0).  A step is one full RK4 step of all 8 fields on
all 8,388,608 cells; the metric is whole-job cell-updates per second with the
state resident in HBM.

Prints ONE JSON line with the throughput, the roofline of the dominant kernel
(timed with HIP events on the library's stream) and a CPU baseline measured on
this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 z·v cell-updates/s + HBM roofline %; surrogate vs solver P_ee sweep speedup"
HBM_PEAK_GBS = 8000.0
VALU_PEAK_GINST = 1024 * 2.4 / 4  # wave64 VALU instructions per ns at the peak clock
# Algorithmic bytes per cell-update (one cell, 8 FP64 fields, one RK4 step):
# the classic accumulator form, per stage read the stage input, y and the
# accumulator and write the accumulator and the next stage input -- stage 1
# 64+128, stages 2 and 3 192+128, stage 4 192+64, 1088 in all.  The roofline
# is defined with these figures: a launch's algorithmic bytes are those of the
# stages it performs.
ALG_STAGE_BYTES = (192, 320, 320, 256)
BYTES_PER_CELL_STEP_DEF = sum(ALG_STAGE_BYTES)
# Bytes per cell each stage kernel actually has to move, by FAST variant:
#   1 (one lane per bin, classic form): 64+128, 192+128, 192+128, 192+64
#   2 (two lanes per bin, stage-value form: no accumulator; stage 4 reads
#      y, Y2, Y3 pointwise): 64+64, 128+64, 128+64, 256+64
#   3 (barrier-free, stage-value form) as 2, plus the per-z partial moments
#      of the input read and of the output written: 2 + 2 bytes per cell
#      (one 64-byte slot per z row and group of 32 bins, each way)
#   4 (two stages per launch; the second stage's input stays in registers):
#      stages 1+2 read y, write Y3 and Q = (Y2 - y) + 2 (Y3 - y) (64+128);
#      stages 3+4 read Y3, y, Q, write the new state (192+64); the timing
#      slots are the two launches
#   5 (FD: four stages per launch; stage outputs stay in registers, the
#      increment sum in LDS): read y, write the new state (64+64), plus the
#      halo rows each team reads again (12 rows per team range)
STAGE_BYTES = {1: (192, 320, 320, 256), 2: (128, 192, 192, 320), 3: (132, 196, 196, 324),
               4: (192, 256, 0, 0), 5: (128, 0, 0, 0)}


def cpu_baseline(scheme: str = "fv", nvz: int = 128, dz: float = 3000.0 / 65536,
                 seconds_target: float = 12.0):
    """Time the reference CPU code (oracle/_ref, built from the reference's own
    headers) or, where that build is absent, the oracle restatement, on a
    z-subsampled slice of the workload: same scheme, dz and nvz, nz = 2048."""
    import numpy as np
    import oracle as O
    from cosenu_amd.initial import seeded_ic

    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    nz = 2048
    p = O.RunParams(nz=nz, nvz=nvz, z0=-nz * dz / 2, z1=nz * dz / 2, flags=O.VAC | O.COLL,
                    omega=0.0, theta=1e-6, scheme=O.FD if scheme == "fd" else O.FV)
    G, bG, s = seeded_ic(p, 0.6, 0.53, 0.92, seed=1)
    use_ref = O.ref_available(O.ref_variant(p))
    if use_ref:
        import tempfile
        run = O.RefRun(p, tempfile.mkdtemp(prefix="cosenu_ref_"))
        fn = lambda st, n: run.step(st, n)
        kind = "reference"
    else:
        fn = lambda st, n: O.step(p, st, n, threads=threads)
        kind = "port"
    # the reference prints a banner per run from C++: keep stdout to the one
    # JSON line by pointing file descriptor 1 at /dev/null meanwhile
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        t0 = time.perf_counter()
        fn(s, 1)
        t1 = time.perf_counter()
        per = max(t1 - t0, 1e-3)
        n = max(1, min(400, int(seconds_target / per)))
        t0 = time.perf_counter()
        fn(s, n)
        el = time.perf_counter() - t0
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    cells = nz * nvz
    return {"value": cells * n / el, "unit": "cell-updates/s", "cores": threads, "kind": kind,
            "sample": f"{n} RK4 steps of {'FD' if scheme == 'fd' else 'FV-WENO7'} (vac+coll) on "
                      f"nz={nz} x nvz={nvz}, same dz as the workload, {threads} OpenMP threads, "
                      f"{el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scheme", choices=["fv", "fd"], default="fv")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c4", "wide"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(the solver would shard as z-slabs with a periodic halo exchange).",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import cosenu_amd as C
    from cosenu_amd.initial import workload

    wl = workload(args.workload, mode=C.FAST, scheme=C.FV if args.scheme == "fv" else C.FD)
    p = wl.params
    a, b, G1, bG1 = wl.factors()  # the separable initial state, built on the device
    cells = p.nz * p.nvz
    sim = C.Simulation(p)
    sim.set_profile_v(G1, bG1)
    sim.init_separable(a, b, C.STATE0)
    sim.init_separable(a, b, C.STATE)

    sim.step(args.warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)   # graph replay, as a caller steps
    sim.sync()
    el = time.perf_counter() - t0
    value = cells * args.steps / el
    ms_step = 1e3 * el / args.steps
    # Per-launch durations: the same number of steps again, with HIP events
    # recorded around every launch on the library's stream.  (The events
    # replace graph replay and add a gap per launch: at the 1024 x 64 size,
    # where a step is ~30 us, they would lengthen it by a third, so the
    # throughput above is timed without them.)
    sim.set_timing(True)
    sim.sync()
    sim.step(args.steps)
    sim.sync()
    ms, cnt = sim.timing()
    sim.set_timing(False)

    # Per-kernel times from HIP events on the library's stream.  Stages 2 and
    # 3 share one kernel; the dominant kernel is the one with the largest
    # total time per step.
    variant = sim.info()["fast_variant"]
    sb = STAGE_BYTES[variant]
    ab = ALG_STAGE_BYTES
    avg = [m / max(1, c) for m, c in zip(ms, cnt)]
    # name: (avg ms per launch, algorithmic bytes/cell, moved bytes/cell, kernel, launches per step)
    if variant == 5:
        kernels = {"stages1-4": (avg[0], sum(ab), sb[0], sim.stage_kernel_name(0), 1)}
    elif variant == 4:
        kernels = {"stages1+2": (avg[0], ab[0] + ab[1], sb[0], sim.stage_kernel_name(0), 1),
                   "stages3+4": (avg[1], ab[2] + ab[3], sb[1], sim.stage_kernel_name(1), 1)}
    else:
        kernels = {"stage1": (avg[0], ab[0], sb[0], sim.stage_kernel_name(0), 1),
                   "stage2+3": ((ms[1] + ms[2]) / max(1, cnt[1] + cnt[2]), ab[1], sb[1],
                                sim.stage_kernel_name(1), 2),
                   "stage4": (avg[3], ab[3], sb[3], sim.stage_kernel_name(3), 1)}
    dom = max(kernels, key=lambda k: kernels[k][0] * kernels[k][4])
    dom_ms, dom_alg, dom_moved, dom_name, _ = kernels[dom]
    achieved = dom_alg * cells / (dom_ms * 1e-3) / 1e9
    step_kernel_ms = sum(avg)
    step_gbs = sum(sb) * cells / (step_kernel_ms * 1e-3) / 1e9
    # HBM bytes per launch of the dominant kernel, from the rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs recorded under profiles/ (scaled by the same counters on a
    # plain copy); used only when they were taken with this launch configuration.
    info = sim.info()
    launch = {"workload": args.workload, "scheme": args.scheme, "fast_variant": variant,
              "z_chunk": info["z_chunk"], "range_teams": info["range_teams"]}
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.scheme}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            rec = json.load(f)
        if rec.get("config") == launch:
            traffic = rec.get(dom)

    # VALU issue rate of the dominant kernel: wavefront VALU instructions per
    # launch from a `rocprofv3 --pmc SQ_INSTS_VALU` run recorded under
    # profiles/ (same launch configuration), over the measured duration,
    # against 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 FP64 instruction.
    valu = None
    vfile = os.path.join(ROOT, "profiles", f"valu_{args.workload}_{args.scheme}.json")
    if os.path.exists(vfile):
        with open(vfile) as f:
            rec = json.load(f)
        if rec.get("config") == launch and dom in rec:
            ginst = rec[dom] / (dom_ms * 1e-3) / 1e9
            valu = {"wave_insts_per_launch": rec[dom], "achieved": ginst, "peak": VALU_PEAK_GINST,
                    "unit": "G wave-instructions/s", "frac": ginst / VALU_PEAK_GINST}

    # Device-copy bandwidth on the same device: one state (nz x 8 x VP doubles)
    # copied between the context's two state buffers (hipMemcpyAsync, read +
    # write), less the per-call synchronisation measured by a same-buffer call.
    sim.copy_state(C.STATE, C.STATE0)
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.copy_state(C.STATE, C.STATE)  # no copy: the call's synchronisation only
    t_sync = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        sim.copy_state(C.STATE, C.STATE0)
    t_copy = (time.perf_counter() - t0) / reps - t_sync
    state_bytes = p.nz * 8 * ((p.nvz + 63) // 64 * 64) * 8
    copy_gbs = 2 * state_bytes / t_copy / 1e9 if t_copy > 0 else None

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {args.scheme.upper()} RK4, nz={p.nz} on "
                               f"[{p.z0:g},{p.z1:g}], nvz={p.nvz}, CFL={p.CFL}, periodic, "
                               f"mu=1, omega=0, FAST mode", "cells": cells,
                   "z_chunk": info["z_chunk"], "range_teams": info["range_teams"],
                   "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dom_name, "kernel_stage": dom,
                     "kernel_alg_bytes_per_cell": dom_alg, "kernel_avg_ms": dom_ms,
                     "kernel_moved_bytes_per_cell": dom_moved,
                     "kernel_moved_GBs": dom_moved * cells / (dom_ms * 1e-3) / 1e9,
                     "fast_variant": variant,
                     "stage_moved_bytes_per_cell": list(sb),
                     "stage_avg_ms": avg,
                     "step_moved_GBs": step_gbs,
                     "frac_at_1088B_per_cell_update": value * BYTES_PER_CELL_STEP_DEF / 1e9 / HBM_PEAK_GBS,
                     "copy_GBs": copy_gbs,
                     "frac_of_copy": achieved / copy_gbs if copy_gbs else None,
                     "moved_frac_of_copy": (dom_moved * cells / (dom_ms * 1e-3) / 1e9 / copy_gbs
                                            if copy_gbs else None),
                     "valu": valu},
    }
    sim.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.scheme, p.nvz, (p.z1 - p.z0) / p.nz)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
