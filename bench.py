#!/usr/bin/env python3
"""Benchmark of the COSEnu time-step hot path on one MI355X.

Headline workload (the single-GPU production configuration): FV-WENO7 + RK4,
FP64, nz = 65536 cells over z in [-1500, 1500], nvz = 128 velocity bins, CFL
0.4, periodic, mu = 1, omega = 0; synthetic flavour-unstable initial state
(sigma_nu = 0.6, sigma_nubar = 0.53, alpha = 0.92, 1e-6 random-phase seed plus
a Gaussian seed at z = 0).  A step is one full RK4 step of all 8 fields on
all 8,388,608 cells; `value` is whole-job cell-updates per second with the
state resident in HBM.

The same JSON line carries
  roofline      the dominant kernel of the headline workload, timed with HIP
                events on the library's stream, against 8 TB/s (algorithmic
                bytes) plus its measured limiter (HBM or FP64 VALU);
  workloads     the other single-GPU workloads: the production grid with the
                other scheme, nz = 2^21 x nvz = 256 (181 GB resident) FV and
                FD, and the wide velocity grid nz = 16384 x nvz = 512 FV and
                FD, each with its dominant kernel;
  batched_sweep the 256-run batch of uneven runs (nz log-uniform in [4096,
                32768], nvz in {64, 128}, random ELN parameters and phases)
                in one graph-captured launch per step and stage group, FV and
                FD: aggregate cell-updates/s and runs checked bit for bit
                against the same runs advanced alone;
  sweep_speedup the surrogate-screened 256-run ELN sweep (nz = 8192, nvz = 64,
                8000 steps) against the batched solver on the same runs;
  cpu_baseline  the reference's own CPU code on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 z·v cell-updates/s + HBM roofline %; surrogate vs solver P_ee sweep speedup"
HBM_PEAK_GBS = 8000.0
SIMDS = 1024            # 256 CUs x 4 SIMDs
PEAK_CLOCK_GHZ = 2.4
# Algorithmic bytes per cell-update (one cell, 8 FP64 fields, one RK4 step):
# the classic accumulator form, per stage read the stage input, y and the
# accumulator and write the accumulator and the next stage input -- stage 1
# 64+128, stages 2 and 3 192+128, stage 4 192+64, 1088 in all.  The roofline
# is defined with these figures: a launch's algorithmic bytes are those of the
# stages it performs.
ALG_STAGE_BYTES = (192, 320, 320, 256)
BYTES_PER_CELL_STEP_DEF = sum(ALG_STAGE_BYTES)
# Bytes per cell each launch actually has to move, by FAST variant:
#   2 (one stage per launch, stage-value form: no accumulator; stage 4 reads
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
#      halo rows each team reads again (12 rows per team range); on rows of
#      512 bins a team spans 2 blocks, which exchange the four stages'
#      moment partials (under 4 bytes per cell)
#   6 (rows of 512 or 1024 bins, two stages per launch with a team over 2 or
#      4 blocks): as 4; the per-row moment partials the blocks exchange
#      (768 bytes per row and block, each way) are under 2 bytes per cell
STAGE_BYTES = {2: (128, 192, 192, 320), 3: (132, 196, 196, 324), 4: (192, 256, 0, 0),
               5: (128, 0, 0, 0), 6: (192, 256, 0, 0)}
# untimed steps before the per-launch event pass (the clock settles over the
# first ~25 ms of sustained work: ~20 steps of the headline workload)
SETTLE_STEPS = 16
SURROGATE_FILE = os.path.join(ROOT, "cosenu_amd", "data", "surrogate_ensemble.npz")
SURROGATE_RECORD = os.path.join(ROOT, "profiles", "surrogate_screened_sweep.json")


def host_cpu():
    """Model name, physical cores and the cores this process may use."""
    model, sockets, cores_per = None, None, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Socket(s)":
                sockets = int(v)
            elif k.strip() == "Core(s) per socket":
                cores_per = int(v)
    except Exception:
        pass
    physical = sockets * cores_per if sockets and cores_per else None
    usable = len(os.sched_getaffinity(0))
    return {"model": model, "physical_cores": physical, "usable_cpus": usable}


def cpu_baseline(scheme: str = "fv", nvz: int = 128, dz: float = 3000.0 / 65536,
                 seconds_target: float = 12.0):
    """Time the reference CPU code (oracle/_ref, built from the reference's own
    headers with its flags) or, where that build is absent, the oracle
    restatement, on a z-subsampled slice of the workload: same scheme, dz and
    nvz, nz = 2048.  OpenMP threads: the CPUs this process may use, at most
    16 (one GPU's share of the host), pinned one per core."""
    cpu = host_cpu()
    threads = max(1, min(16, cpu["usable_cpus"]))
    # libgomp reads these when the oracle libraries are first loaded
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    import oracle as O
    from cosenu_amd.initial import seeded_ic

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
            "host": cpu, "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close",
                                 "OMP_PLACES": "cores"},
            "sample": f"{n} RK4 steps of {'FD' if scheme == 'fd' else 'FV-WENO7'} (vac+coll) on "
                      f"nz={nz} x nvz={nvz}, same dz as the workload, {threads} OpenMP threads "
                      f"(one per core), {el:.1f} s"}


def _profile_record(kind, workload_name, scheme, launch):
    """A committed rocprofv3 record (profiles/<kind>_<workload>_<scheme>.json)
    taken with this launch configuration, or None."""
    f = os.path.join(ROOT, "profiles", f"{kind}_{workload_name}_{scheme}.json")
    if not os.path.exists(f):
        return None
    with open(f) as fh:
        rec = json.load(fh)
    return rec if rec.get("config") == launch else None


def measure(C, name, scheme, steps, warmup, timing_steps):
    """Per-launch durations from HIP events over `timing_steps` steps (after
    max(warmup, SETTLE_STEPS) untimed steps), then the throughput of `steps`
    RK4 steps after `warmup` more (graph replay, the graphs captured before
    anything is timed).  Returns (result dict, open Simulation)."""
    from cosenu_amd.initial import workload
    wl = workload(name, mode=C.FAST, scheme=C.FV if scheme == "fv" else C.FD)
    p = wl.params
    a, b, G1, bG1 = wl.factors()  # the separable initial state, built on the device
    cells = p.nz * p.nvz
    sim = C.Simulation(p)
    sim.set_profile_v(G1, bG1)
    sim.init_separable(a, b, C.STATE0)
    sim.init_separable(a, b, C.STATE)

    # the step graphs (one per state buffer the step starts from) are captured
    # before anything is timed
    sim.prepare()
    # The chip's clock rises over the first ~25 ms of sustained work after
    # the GPU was idle: neither measurement below starts before SETTLE_STEPS
    # untimed steps (or the warm-up count, if larger) have run.
    sim.step(max(warmup, SETTLE_STEPS))
    sim.sync()
    # Per-launch durations: HIP events recorded around every launch on the
    # library's stream, over `timing_steps` steps.  (The events replace
    # graph replay and add a gap per launch: at the 1024 x 64 size, where a
    # step is ~30 us, they would lengthen it by a third, so the throughput
    # below is timed without them.)  The warm-up steps and the timed window
    # follow, so the throughput is that of the clock the chip holds under
    # sustained load, as a long run sees it.
    sim.set_timing(True)
    sim.sync()
    sim.step(timing_steps)
    sim.sync()
    ms, cnt = sim.timing()
    sim.set_timing(False)
    sim.step(warmup)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps)   # graph replay, as a caller steps
    sim.sync()
    el = time.perf_counter() - t0
    d = sim.diagnostics()

    info = sim.info()
    variant = info["fast_variant"]
    sb = STAGE_BYTES[variant]
    ab = ALG_STAGE_BYTES
    avg = [m / max(1, c) for m, c in zip(ms, cnt)]
    # name: (avg ms per launch, algorithmic bytes/cell, moved bytes/cell, kernel, launches per step)
    if variant == 5:
        kernels = {"stages1-4": (avg[0], sum(ab), sb[0], sim.stage_kernel_name(0), 1)}
    elif variant in (4, 6):
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
    moved_gbs = dom_moved * cells / (dom_ms * 1e-3) / 1e9
    launch = {"workload": name, "scheme": scheme, "fast_variant": variant,
              "z_chunk": info["z_chunk"], "range_teams": info["range_teams"]}

    # FP64 VALU issue of the dominant kernel: wavefront VALU instructions per
    # launch and the clock the chip held, from a committed `rocprofv3 --pmc
    # SQ_INSTS_VALU ... GRBM_GUI_ACTIVE` record of this launch configuration,
    # over the duration measured here; one wave64 FP64 instruction per SIMD
    # every 4 cycles is the issue ceiling.
    valu = None
    rec = _profile_record("valu", name, scheme, launch)
    if rec and dom in rec:
        ginst = rec[dom] / (dom_ms * 1e-3) / 1e9
        clock = (rec.get(dom + "_detail") or {}).get("clock_GHz")
        valu = {"wave_insts_per_launch": rec[dom], "achieved": ginst,
                "peak_at_2.4GHz": SIMDS * PEAK_CLOCK_GHZ / 4, "unit": "G wave-instructions/s",
                "frac_at_2.4GHz": ginst / (SIMDS * PEAK_CLOCK_GHZ / 4)}
        if clock:
            valu.update({"measured_clock_GHz": clock,
                         "frac_at_measured_clock": ginst / (SIMDS * clock / 4),
                         "valu_busy": rec[dom + "_detail"].get("valu_busy")})
    traffic_rec = _profile_record("traffic", name, scheme, launch)
    traffic = traffic_rec.get(dom) if traffic_rec else None
    res = {
        "workload": f"{wl.name}: {scheme.upper()} RK4, nz={p.nz} on [{p.z0:g},{p.z1:g}], "
                    f"nvz={p.nvz}, CFL={p.CFL}, {'open' if p.bc else 'periodic'}, mu=1, omega=0, "
                    "FAST mode",
        "cells": cells, "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * el / steps,
        "value": cells * steps / el, "unit": "cell-updates/s",
        "frac_at_1088B_per_cell_update": cells * steps / el * BYTES_PER_CELL_STEP_DEF / 1e9
        / HBM_PEAK_GBS,
        "fast_variant": variant, "z_chunk": info["z_chunk"], "range_teams": info["range_teams"],
        "device_GB": info["device_bytes"] / 1e9,
        "kernel": dom_name, "kernel_stage": dom, "kernel_avg_ms": dom_ms,
        "kernel_alg_bytes_per_cell": dom_alg, "kernel_alg_GBs": achieved,
        "kernel_frac": achieved / HBM_PEAK_GBS,
        "kernel_moved_bytes_per_cell": dom_moved, "kernel_moved_GBs": moved_gbs,
        "stage_avg_ms": avg, "stage_moved_bytes_per_cell": list(sb),
        "timing_launches": cnt, "traffic": traffic, "valu": valu,
        "nonfinite": d["nonfinite"],
    }
    return res, sim


def copy_rate(C, sim, p):
    """Device-copy bandwidth: one state (nz x 8 x VP doubles) copied between the
    context's two state buffers (hipMemcpyAsync, read + write bytes), less the
    per-call synchronisation measured by a same-buffer call."""
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
    vp = 64 if p.nvz <= 64 else 128 if p.nvz <= 128 else 256 if p.nvz <= 256 else \
        (p.nvz + 63) // 64 * 64
    state_bytes = p.nz * 8 * vp * 8
    return 2 * state_bytes / t_copy / 1e9 if t_copy > 0 else None


def limiter(res, copy_gbs):
    """The dominant kernel's measured limiter: the larger of its HBM use (bytes
    it moves per second against the device-copy rate) and its FP64 VALU issue
    (against the ceiling at the clock the chip held)."""
    hbm = res["kernel_moved_GBs"] / copy_gbs if copy_gbs else None
    v = res["valu"]
    valu = v.get("frac_at_measured_clock", v["frac_at_2.4GHz"]) if v else None
    if valu is not None and (hbm is None or valu > hbm):
        return "valu", hbm, valu
    return "hbm", hbm, valu


def sweep_speedup(C, window: int, log):
    """The surrogate-screened 256-run ELN sweep against the batched solver.

    Solver: the 256 runs (nz = 8192 on [-600, 600], nvz = 64, 8000 steps, P_ee(v)
    averaged over the final 20% of the steps, sampled every 20) in one batch;
    `window` batch steps are timed after a graph-capturing warm-up and the
    8000-step time is extrapolated linearly, plus the 80 samples of the
    average (one timed).  Screened, for each of the ensemble's two gates
    (sweep-mean and per-run): features, ensemble inference on the device,
    then the rejected runs alone through the same solver schedule (timed the
    same way).  The accuracy of the screened results is the committed record
    of the full pipeline run, which solved all 256 runs."""
    import numpy as np
    from cosenu_amd import surrogate as S
    if not os.path.exists(SURROGATE_FILE):
        return {"status": "no trained ensemble", "file": os.path.relpath(SURROGATE_FILE, ROOT)}
    setup = S.SolverSetup(nz=8192, steps=8000)
    surr = S.Surrogate.load(SURROGATE_FILE)
    elns = S.sample_elns(256, seed=surr.meta.get("sweep_seed", 9001))
    p = S.run_params(setup)
    nsamp = len(range(int(round(setup.steps * (1 - setup.avg_frac))), setup.steps + 1,
                      setup.sample_every))

    def solver_seconds(chunk):
        """Extrapolated seconds of the solver schedule for these runs."""
        with C.Batch([p] * len(chunk)) as B:
            for r, e in enumerate(chunk):
                G, bG, s0 = S.seeded_ic(p, e.sig_nu, e.sig_nubar, e.alpha, e.seed)
                B.set_profile(r, G[:, :1], bG[:, :1])
                B.upload(r, s0, C.STATE0)
                B.upload(r, s0, C.STATE)
            B.step(16)  # captures the 16-step graph
            B.sync()
            t0 = time.perf_counter()
            B.step(window)
            B.sync()
            per_step = (time.perf_counter() - t0) / window
            t0 = time.perf_counter()
            for r in range(len(chunk)):
                B.survival_profile(r)
            per_sample = time.perf_counter() - t0
        return setup.steps * per_step + nsamp * per_sample, per_step, per_sample

    log("sweep: solver-only timing")
    t_solver, step_all, samp_all = solver_seconds(elns)
    out = {"runs": 256, "grid": "nz=8192 on [-600,600], nvz=64, 8000 steps, FV, P_ee(v) "
                                 "averaged over steps 6400-8000 every 20",
           "window_steps": window, "solver_s_per_step": step_all,
           "solver_s_per_sample": samp_all, "solver_only_s": t_solver,
           "members": len(surr.members), "bound": surr.bound}
    t0 = time.perf_counter()
    X = np.stack([S.features(e) for e in elns])
    t_feat = time.perf_counter() - t0
    surr.predict(X[:8])  # first call loads the kernels
    for gate, key in (("mean", "sweep_mean_gate"), ("run", "per_run_gate")):
        log(f"sweep: screened timing, {gate} gate")
        t0 = time.perf_counter()
        _, accept = surr.screen(X, gate=gate)
        t_inf = time.perf_counter() - t0
        rejected = [e for e, ok in zip(elns, accept) if not ok]
        t_fb = solver_seconds(rejected)[0] if rejected else 0.0
        t_screened = t_feat + t_inf + t_fb
        out[key] = {"tau": surr.tau if gate == "run" else surr.tau_mean,
                    "features_s": t_feat, "inference_s": t_inf, "fallback_runs": len(rejected),
                    "fallback_s": t_fb, "screened_s": t_screened,
                    "speedup": t_solver / t_screened}
    # the headline figure: the gate that holds the bound on the sweep's mean
    out["speedup"] = out["sweep_mean_gate"]["speedup"]
    if os.path.exists(SURROGATE_RECORD):
        with open(SURROGATE_RECORD) as fh:
            rec = json.load(fh)
        out["accuracy_from_pipeline_record"] = {
            k: rec.get(k) for k in ("per_run_gate", "sweep_mean_gate", "heldout", "bound",
                                    "solver_seed_reproducibility", "dataset")}
    for m in surr.members:
        m.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scheme", choices=["fv", "fd"], default="fv")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", action="store_true",
                    help="the headline workload only (no other workloads, no sweep)")
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c1_open", "c4", "wide", "wide_open"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: every GPU box here has one, and the library "
              "has no multi-GPU path (the solver would shard as z-slabs with a periodic halo "
              "exchange).", file=sys.stderr)
        sys.exit(2)

    import cosenu_amd as C

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    log(f"headline: {args.workload} {args.scheme}, {args.steps} steps after {args.warmup}")
    res, sim = measure(C, args.workload, args.scheme, args.steps, args.warmup,
                       timing_steps=args.steps)
    copy_gbs = copy_rate(C, sim, sim.p)
    sim.close()
    bound, hbm_frac, valu_frac = limiter(res, copy_gbs)

    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": res["workload"], "cells": res["cells"],
                   "z_chunk": res["z_chunk"], "range_teams": res["range_teams"],
                   "parallelism": "1 GPU"},
        "roofline": {"bound": bound, "achieved": res["kernel_alg_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": res["kernel_frac"], "traffic": res["traffic"],
                     "kernel": res["kernel"], "kernel_stage": res["kernel_stage"],
                     "kernel_alg_bytes_per_cell": res["kernel_alg_bytes_per_cell"],
                     "kernel_avg_ms": res["kernel_avg_ms"],
                     "kernel_moved_bytes_per_cell": res["kernel_moved_bytes_per_cell"],
                     "kernel_moved_GBs": res["kernel_moved_GBs"],
                     "fast_variant": res["fast_variant"],
                     "stage_moved_bytes_per_cell": res["stage_moved_bytes_per_cell"],
                     "stage_avg_ms": res["stage_avg_ms"],
                     "frac_at_1088B_per_cell_update": res["frac_at_1088B_per_cell_update"],
                     "copy_GBs": copy_gbs,
                     "frac_of_copy": res["kernel_alg_GBs"] / copy_gbs if copy_gbs else None,
                     "hbm_use_of_copy": hbm_frac,
                     "valu_use_at_measured_clock": valu_frac,
                     "valu": res["valu"]},
    }
    if not args.only:
        # the other workloads: K steps (at most 40) after 16 warm-up steps,
        # one graph replay, so that their short timed region sees a settled
        # clock (20 steps after 2: the c1 FD step read 0.436 ms against
        # 0.395 ms of kernel time); each also settles before its event pass
        side, side_warm = min(args.steps, 40), 16
        others = [("c1", "fd" if args.scheme == "fv" else "fv"), ("c4", "fv"), ("c4", "fd"),
                  ("wide", "fv"), ("wide", "fd")]
        out["workloads"] = {}
        for name, scheme in others:
            if (name, scheme) == (args.workload, args.scheme):
                continue
            log(f"workload {name} {scheme}: {side} steps after {side_warm}")
            r, s2 = measure(C, name, scheme, side, side_warm, timing_steps=side)
            s2.close()
            r["bound"], r["hbm_use_of_copy"], r["valu_use_at_measured_clock"] = limiter(r, copy_gbs)
            out["workloads"][f"{name}_{scheme}"] = r
        from cosenu_amd.sweep import run_sweep
        out["batched_sweep"] = {}
        for scheme in ("fv", "fd"):
            log(f"batched sweep {scheme}: 256 runs, {side} steps after {side_warm}")
            out["batched_sweep"][scheme] = run_sweep(256, steps=side, warmup=side_warm, check=2,
                                                     scheme=scheme)
        out["sweep_speedup"] = sweep_speedup(C, window=160, log=log)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.scheme, sim.p.nvz, sim.p.dz)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
