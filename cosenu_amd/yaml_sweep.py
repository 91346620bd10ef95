"""Sweep front-end: the reference's cosenu_configs.yaml, run as one batch.

    python -m cosenu_amd.yaml_sweep cosenu_configs.yaml --scheme fv --out DIR [--run]

Expands the YAML's nzs x nvzs x CFLS into one job per combination exactly as
the reference's `manage.py configure()` does (manage.py:142-282): folder
`<out>/<nz>_<nvz>_<CFL>/` with a `job.config` (gz = 4 for FV, 2 for FD;
dz = (z1-z0)/nz; dt = |CFL dz / v1|; N_ITER = int(end_time/dt);
ANAL_EVERY = max(1, int(N_ITER/n_analyze)); the same keys and `key : value`
layout), plus `jobs_list.txt`.  The compile-time presets the reference writes
into presets.hpp become runtime keys appended to each job.config (the
reference's parser reports them as unknown and ignores them, so the files
stay usable by the reference too).

With --run, all jobs advance together in one cosenu_batch (FAST mode,
periodic boundaries) from the reference's initial condition
(initialize.hpp:6-47 with the YAML's perturbation_size), and each job folder
receives the reference's output files at that job's own cadence:
conserved_quantities.dat, survival_probability.dat, rho_<t>.dat (every
N_ITER/n_dump_rho steps when n_dump_rho > 0) and time.txt.  Jobs that need
STRICT mode or open boundaries run one by one with `cosenu_amd/bin/cosenu
--conf job.config` instead (--driver).
"""
from __future__ import annotations

import argparse
import math
import os
import subprocess
import sys
import time

import numpy as np
import yaml

from . import (ADVEC_OFF, COLL, FAST, FD, FV, MAT, OPEN, PERIODIC, STATE, STATE0, VAC, Batch,
               RunParams)
from .initial import g_profile


def write_dict(d: dict, path: str):
    """The reference's job.config layout (manage.py:116-121)."""
    maxlen = max(len(k) for k in d)
    with open(path, "w") as f:
        for k, v in d.items():
            f.write(f"{str(k).ljust(maxlen + 1)} : {str(v)} \n")


def expand(cfg: dict, scheme: str) -> list:
    """Jobs of a cosenu_configs.yaml, as manage.py configure() derives them."""
    z = cfg["zrange"]
    v0, v1 = cfg["v0"], cfg["v1"]
    gz = 2 if scheme == "fd" else 4
    jobs = []
    for nz in cfg["nzs"]:
        for nvz in cfg["nvzs"]:
            for CFL in cfg["CFLS"]:
                z0, z1 = z[0], z[1]
                dz = (z1 - z0) / nz
                dt = abs(CFL * dz / v1)
                N_ITER = int(cfg["end_time"] / dt)
                nan = cfg["n_analyze"]
                ANAL_EVERY = max(1, int(N_ITER / nan)) if nan != 0 else N_ITER + 1
                job = {"scheme": scheme.upper(), "z0": z[0], "z1": z[1], "nvz": nvz, "CFL": CFL,
                       "dz": dz, "dt": dt, "gz": gz, "v0": v0, "v1": v1, "nz": nz,
                       "pmo": cfg["pmo"], "omega": cfg["omega"], "theta": cfg["theta"],
                       "mu": cfg["mu"], "perturbation_size": cfg["perturbation_size"],
                       "N_ITER": N_ITER, "ANAL_EVERY": ANAL_EVERY,
                       "n_dump_rho": cfg["n_dump_rho"]}
                jobs.append((f"{nz}_{nvz}_{CFL}", job))
    return jobs


def presets(cfg: dict) -> dict:
    return {"boundary": "open" if cfg.get("boundary") == "open" else "periodic",
            "vac_osc_on": bool(cfg.get("vac_osc_on", True)),
            "mat_osc_on": bool(cfg.get("mat_osc_on", False)),
            "collective_osc_on": bool(cfg.get("collective_osc_on", True)),
            "advection_off": bool(cfg.get("advection_off", False))}


def run_params(job: dict, pre: dict, mode: int = FAST) -> RunParams:
    flags = (VAC if pre["vac_osc_on"] else 0) | (MAT if pre["mat_osc_on"] else 0) | \
            (COLL if pre["collective_osc_on"] else 0) | (ADVEC_OFF if pre["advection_off"] else 0)
    return RunParams(nz=int(job["nz"]), nvz=int(job["nvz"]), z0=float(job["z0"]),
                     z1=float(job["z1"]), v0=float(job["v0"]), v1=float(job["v1"]),
                     CFL=float(job["CFL"]), scheme=FD if job["scheme"] == "FD" else FV,
                     bc=OPEN if pre["boundary"] == "open" else PERIODIC, flags=flags,
                     pmo=float(job["pmo"]), omega=float(job["omega"]), theta=float(job["theta"]),
                     mu=float(job["mu"]), gz=int(job["gz"]), mode=mode)


def reference_factors(p: RunParams, perturbation_size: float, signu=0.6, sigbnu=0.6, alpha=0.9):
    """NuOsc::initialize in separable form: field f at (v, z) = a[f, v] * b[f, z]
    with the reference's own products (0.5 G(v)) (1 +- sqrt(1 - eps(z)^2)), ..."""
    dv = (p.v1 - p.v0) / p.nvz
    dz = (p.z1 - p.z0) / p.nz
    vz = [p.v0 + (i + 0.5) * dv for i in range(p.nvz)]
    Z = [p.z0 + (j + 0.5) * dz for j in range(p.nz)]
    G = np.array([g_profile(v, 1.0, signu) for v in vz])
    bG = np.array([alpha * g_profile(v, 1.0, sigbnu) for v in vz])
    e = np.array([perturbation_size * math.exp(-(z - 0.0) * (z - 0.0) / 50.0) for z in Z])
    e_ = np.array([math.sqrt(1.0 - x * x) for x in e])
    one = np.ones(p.nvz)
    a = np.stack([0.5 * G, 0.5 * G, 0.5 * G, one, 0.5 * bG, 0.5 * bG, 0.5 * bG, one])
    b = np.stack([1.0 + e_, 1.0 - e_, 0.0 + e, np.full(p.nz, -0.0),
                  1.0 + e_, 1.0 - e_, 0.0 + e, np.zeros(p.nz)])
    return a, b, G, bG


def _dcon_line(t, d):
    return (f"{t}\t{d['dP_max']:.20f}\t{d['dP_avg']:.20f}\t{d['dbP_avg']:.20f}\t"
            f"{d['M0_norm']:.20f}\n")


def _dump_rho(path, p: RunParams, state: np.ndarray):
    dv = (p.v1 - p.v0) / p.nvz
    dz = (p.z1 - p.z0) / p.nz
    vz = [p.v0 + (i + 0.5) * dv for i in range(p.nvz)]
    Z = [p.z0 + (j + 0.5) * dz for j in range(p.nz)]
    first = True
    with open(path, "w") as f:
        for j in range(p.nz):
            for i in range(p.nvz):
                vals = "\t".join(format(float(state[k, i, j]), ".16g") for k in range(8))
                if first:  # the reference sets precision 16 after the first z and v
                    f.write(f"{format(Z[j], 'g')}\t{format(vz[i], 'g')}\t{vals}\n")
                    first = False
                else:
                    f.write(f"{format(Z[j], '.16g')}\t{format(vz[i], '.16g')}\t{vals}\n")
            f.write("\n")


def run_batch(jobs, pre, out, log=print):
    runs = [run_params(job, pre) for _, job in jobs]
    dirs = [os.path.join(out, jid) for jid, _ in jobs]
    n_iter = [int(job["N_ITER"]) for _, job in jobs]
    anal = [int(job["ANAL_EVERY"]) for _, job in jobs]
    dumpi = []
    for _, job in jobs:
        nd = int(job["n_dump_rho"])
        dumpi.append(max(1, int(int(job["N_ITER"]) / nd)) if nd > 0 else int(job["N_ITER"]) + 1)
    B = Batch(runs)
    for r, ((_, job), p) in enumerate(zip(jobs, runs)):
        a, b, G, bG = reference_factors(p, float(job["perturbation_size"]))
        st = a[:, :, None] * b[:, None, :]
        B.set_profile(r, G[:, None], bG[:, None])
        B.upload(r, st, STATE0)
        B.upload(r, st, STATE)
    # t = 0 outputs (main.cpp:159-169)
    for r, d in enumerate(dirs):
        dg = B.diagnostics(r)
        with open(os.path.join(d, "conserved_quantities.dat"), "w") as f:
            f.write("# [time, dP_max(t), <dP>(t), <dbP(t), <dP>(t), |M0|]\n" + _dcon_line(0, dg))
        with open(os.path.join(d, "survival_probability.dat"), "w") as f:
            f.write("# [time, <Pee>, <Pbee>]\n" + f"0\t{dg['Pee']:e}\t{dg['Pbee']:e}\n")
        _dump_rho(os.path.join(d, "rho_0.dat"), runs[r], B.download(r))
    t0 = time.perf_counter()
    t = 1
    last = max(n_iter) - 1
    while t <= last:
        # next iteration at which some job records something
        te = t
        while te < last and not any(te < n_iter[r] and (te % anal[r] == 0 or te % dumpi[r] == 0)
                                    for r in range(len(runs))):
            te += 1
        B.step(te - t + 1)
        t = te
        for r, d in enumerate(dirs):
            if t >= n_iter[r]:
                continue
            if t % anal[r] == 0:
                dg = B.diagnostics(r)
                with open(os.path.join(d, "conserved_quantities.dat"), "a") as f:
                    f.write(_dcon_line(t, dg))
                with open(os.path.join(d, "survival_probability.dat"), "a") as f:
                    f.write(f"{t}\t{dg['Pee']:e}\t{dg['Pbee']:e}\n")
            if t % dumpi[r] == 0:
                _dump_rho(os.path.join(d, f"rho_{t}.dat"), runs[r], B.download(r))
        t += 1
    B.sync()
    el = time.perf_counter() - t0
    for d in dirs:
        with open(os.path.join(d, "time.txt"), "w") as f:
            f.write(f"{el}\n")
    B.close()
    log(f"{len(runs)} jobs, {last} steps in {el:.2f} s")
    return el


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config")
    ap.add_argument("--scheme", choices=["fv", "fd"], default="fv")
    ap.add_argument("--out", default="")
    ap.add_argument("--run", action="store_true", help="run all jobs as one batch (FAST)")
    ap.add_argument("--driver", action="store_true", help="run the jobs one by one with bin/cosenu")
    args = ap.parse_args(argv)
    with open(args.config) as f:
        cfg = yaml.safe_load(f)
    out = args.out or os.path.abspath(cfg.get(f"folder_{args.scheme}", f"folder_{args.scheme}"))
    os.makedirs(out, exist_ok=True)
    jobs = expand(cfg, args.scheme)
    pre = presets(cfg)
    with open(os.path.join(out, "jobs_list.txt"), "w") as jl:
        for jid, job in jobs:
            d = os.path.join(out, jid)
            os.makedirs(d, exist_ok=True)
            write_dict(job, os.path.join(d, "job.config"))
            with open(os.path.join(d, "job.config"), "a") as f:
                for k, v in pre.items():
                    f.write(f"{k} : {v}\n")
            jl.write(f"{jid}, job.config \n")
    print(f"[ OK ]...{len(jobs)} jobs under {out}")
    if args.run:
        if pre["boundary"] == "open":
            print("[ FAIL ]...batched runs take periodic boundaries; use --driver")
            return 1
        run_batch(jobs, pre, out)
    elif args.driver:
        exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "cosenu")
        for jid, _ in jobs:
            r = subprocess.run([exe, "--conf", "job.config"], cwd=os.path.join(out, jid))
            if r.returncode:
                return r.returncode
    return 0


if __name__ == "__main__":
    sys.exit(main())
