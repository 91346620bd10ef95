"""The benchmark workloads at their own sizes, through the C ABI.

* Production grid: nz = 65536 on [-1500, 1500], nvz = 128, FV (two RK4
  stages per launch) and FD (four per launch), the default launch (one
  contiguous z range per resident team).  Three FAST steps against the
  oracle's moment form, one STRICT step bit for bit against the oracle's
  reference arithmetic, and 200 steps FAST against STRICT (STRICT is the
  reference's arithmetic bit for bit: tests/test_gpu_parity.py,
  tests/test_gpu_strict_wide.py), compared as P_ee = ee/ee0, the state, the
  total trace and the lepton number.
* HBM-filling run: nz = 2^21 on [-20000, 20000], nvz = 256 (181 GB on the
  device): conservation of trace, lepton number and M0 at omega = 0 over 12
  steps, no non-finite values, one asynchronous P_ee snapshot checked against
  the device's survival profile, and the same physics on a z-subsampled
  periodic domain (nz = 8192, same dz) against the oracle.
* 256-run sweep: nz log-uniform in [4096, 32768], nvz in {64, 128}, random
  sigma_nu, sigma_nubar, alpha and phases, all in one batch: every run bit
  for bit equal to the same run in its own context, two runs against the
  oracle.
* Surrogate dataset grid: nz = 8192, nvz = 64, 8000 steps, P_ee(v) averaged
  over the final 20% of the steps, batched against a run of its own.

Sizes follow the workloads; each test prints what it checks before the
device work starts.
"""
import dataclasses
import math

import numpy as np
import pytest

import oracle as O
import cosenu_amd as C
from cosenu_amd.initial import workload, seeded_ic
from cosenu_amd.validate_long import subsampled

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _orc(p):
    return O.RunParams(nz=p.nz, nvz=p.nvz, z0=p.z0, z1=p.z1, v0=p.v0, v1=p.v1, CFL=p.CFL,
                       scheme=p.scheme, bc=p.bc, flags=p.flags, pmo=p.pmo, omega=p.omega,
                       theta=p.theta, mu=p.mu, ko=p.ko)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _log(msg):
    print(msg, flush=True)


def _start(p, a, b, G1, bG1):
    sim = C.Simulation(p)
    sim.set_profile_v(G1, bG1)
    sim.init_separable(a, b, C.STATE0)
    sim.init_separable(a, b, C.STATE)
    return sim


@pytest.mark.parametrize("scheme", [C.FV, C.FD])
def test_production_grid(gpu, scheme):
    name = "FV" if scheme == C.FV else "FD"
    wl = workload("c1", mode=C.FAST, scheme=scheme)
    p = wl.params
    po = _orc(p)
    a, b, G1, bG1 = wl.factors()
    s0 = a[:, :, None] * b[:, None, :]
    _log(f"production grid {name}: nz={p.nz} nvz={p.nvz}, FAST 3 + 197 steps")
    with _start(p, a, b, G1, bG1) as sim:
        info = sim.info()
        assert info["fast_variant"] == (4 if scheme == C.FV else 5)
        assert info["range_teams"] > 0
        sim.step(3)
        f3 = sim.download()
        sim.step(197)
        f200 = sim.download()
        dF = sim.diagnostics()
    _log(f"production grid {name}: STRICT 1 + 199 steps")
    with _start(dataclasses.replace(p, mode=C.STRICT), a, b, G1, bG1) as sim:
        sim.step(1)
        s1 = sim.download()
        sim.step(199)
        s200 = sim.download()
        dS = sim.diagnostics()
    _log(f"production grid {name}: oracle, moment form 3 steps, reference arithmetic 1 step")
    ref3 = O.step(po, s0, 3, moment_form=True)
    ref1 = O.step(po, s0, 1)
    r3, r3_2 = _rel(f3, ref3), _rel2(f3, ref3)
    p200 = _rel(f200[0] / s0[0], s200[0] / s0[0])
    p200_2 = _rel2(f200[0] / s0[0], s200[0] / s0[0])
    st200 = _rel(f200, s200)
    tr = abs(dF["trace"] - dS["trace"]) / abs(dS["trace"])
    lp = abs(dF["lepton"] - dS["lepton"]) / abs(dS["lepton"])
    _log(f"production grid {name}: FAST-oracle after 3 steps {r3:.3e} (L2 {r3_2:.3e}); STRICT "
         f"1 step equal to the oracle: {np.array_equal(s1, ref1)}; FAST-STRICT after 200 steps: "
         f"P_ee {p200:.3e} (L2 {p200_2:.3e}), state {st200:.3e}, trace {tr:.3e}, lepton {lp:.3e}")
    assert r3 <= 1e-10 and r3_2 <= 1e-10
    assert np.array_equal(s1, ref1)
    assert p200 <= 1e-10 and p200_2 <= 1e-10 and st200 <= 1e-10
    assert tr <= 1e-14 and lp <= 1e-12
    assert dF["nonfinite"] == 0 and dS["nonfinite"] == 0
    assert not np.array_equal(f200, s0)


@pytest.mark.parametrize("name,scheme,variant", [("c1_open", C.FV, 4), ("c1_open", C.FD, 5),
                                                  ("wide_open", C.FV, 3), ("wide_open", C.FD, 3),
                                                  ("wide_open", C.FV, 6), ("wide_open", C.FD, 6),
                                                  ("wide_open", C.FD, 5)])
def test_open_boundary_full_size(gpu, monkeypatch, name, scheme, variant):
    """The open boundary at full size: nz = 65536 x 128 (two- and four-stage
    kernels) and nz = 16384 x 480 (a bin group across v = 0 and padded rows:
    the general barrier-free kernel and the multi-block two- and four-stage
    kernels).
    FAST against the oracle (moment form, 3 steps) and against STRICT (50
    steps); STRICT's first step equal to the oracle's reference arithmetic."""
    monkeypatch.setenv("COSENU_FAST_VARIANT", str(variant))
    tag = f"{name} {'FV' if scheme == C.FV else 'FD'} variant {variant}"
    wl = workload(name, mode=C.FAST, scheme=scheme)
    p = wl.params
    po = _orc(p)
    a, b, G1, bG1 = wl.factors()
    s0 = a[:, :, None] * b[:, None, :]
    _log(f"{tag}: nz={p.nz} nvz={p.nvz} open, FAST 3 + 47 steps")
    with _start(p, a, b, G1, bG1) as sim:
        assert sim.info()["fast_variant"] == variant
        sim.step(3)
        f3 = sim.download()
        sim.step(47)
        f50 = sim.download()
    _log(f"{tag}: STRICT 1 + 49 steps")
    with _start(dataclasses.replace(p, mode=C.STRICT), a, b, G1, bG1) as sim:
        sim.step(1)
        s1 = sim.download()
        sim.step(49)
        s50 = sim.download()
    _log(f"{tag}: oracle, moment form 3 steps, reference arithmetic 1 step")
    ref3 = O.step(po, s0, 3, moment_form=True)
    ref1 = O.step(po, s0, 1)
    r3, st50 = _rel(f3, ref3), _rel(f50, s50)
    _log(f"{tag}: FAST-oracle after 3 steps {r3:.3e}; STRICT 1 step equal to the oracle: "
         f"{np.array_equal(s1, ref1)}; FAST-STRICT after 50 steps {st50:.3e}")
    assert r3 <= 1e-10 and st50 <= 1e-10
    assert np.array_equal(s1, ref1)
    assert np.isfinite(f50).all() and not np.array_equal(f50, s0)


@pytest.mark.parametrize("scheme,variant", [(C.FV, "6"), (C.FD, "5"), (C.FD, "6")])
def test_wide_grid_full_size(gpu, monkeypatch, scheme, variant):
    """nz = 16384 x 512 at full size (the `wide` workload), periodic: the
    multi-block kernels (four stages per launch, variant 5, the default for
    FD on rows of 512 bins; two stages per launch, variant 6, the default
    for FV) against the oracle (moment form, 3 steps) and against the
    barrier-free kernel (variant 3, 50 steps); total trace conserved."""
    tag = f"wide {'FV' if scheme == C.FV else 'FD'} variant {variant}"
    wl = workload("wide", mode=C.FAST, scheme=scheme)
    p = wl.params
    po = _orc(p)
    a, b, G1, bG1 = wl.factors()
    s0 = a[:, :, None] * b[:, None, :]
    outs = {}
    for v in (variant, "3"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", v)
        _log(f"{tag}: nz={p.nz} nvz={p.nvz} periodic, variant {v}, 3 + 47 steps")
        with _start(p, a, b, G1, bG1) as sim:
            assert sim.info()["fast_variant"] == int(v)
            d0 = sim.diagnostics()
            sim.step(3)
            f3 = sim.download()
            sim.step(47)
            f50 = sim.download()
            d50 = sim.diagnostics()
        outs[v] = (f3, f50, d0, d50)
    _log(f"{tag}: oracle, moment form 3 steps")
    ref3 = O.step(po, s0, 3, moment_form=True)
    f3, f50, d0, d50 = outs[variant]
    r3, r50 = _rel(f3, ref3), _rel(f50, outs["3"][1])
    tr = abs(d50["trace"] - d0["trace"]) / abs(d0["trace"])
    _log(f"{tag}: oracle after 3 steps {r3:.3e}; against variant 3 after 50 "
         f"steps {r50:.3e}; trace drift {tr:.3e}")
    assert r3 <= 1e-10 and r50 <= 1e-12 and tr <= 1e-14
    assert d50["nonfinite"] == 0 and not np.array_equal(f50, s0)


def _drift(d, d0, key):
    return abs(d[key] - d0[key]) / abs(d0[key])


@pytest.mark.parametrize("scheme", [C.FV, C.FD])
def test_hbm_filling_run(gpu, scheme):
    name = "FV" if scheme == C.FV else "FD"
    wl = workload("c4", mode=C.FAST, scheme=scheme)
    p = wl.params
    a, b, G1, bG1 = wl.factors()
    cells = p.nz * p.nvz
    _log(f"HBM-filling run {name}: nz={p.nz} nvz={p.nvz}, 10 + 2 steps, one P_ee snapshot")
    buf = C.PinnedBuffer(cells)
    try:
        with _start(p, a, b, G1, bG1) as sim:
            info = sim.info()
            _log(f"HBM-filling run {name}: {info['device_bytes'] / 1e9:.1f} GB on the device, "
                 f"variant {info['fast_variant']}")
            assert info["device_bytes"] > 170e9
            d0 = sim.diagnostics()
            sim.step(10)
            d10 = sim.diagnostics()
            pee, _ = sim.survival_profile()
            sim.snapshot_pee_async(buf)
            sim.step(2)  # keeps stepping while the snapshot is copied
            sim.snapshot_wait()
            d12 = sim.diagnostics()
        snap = buf.array.reshape(p.nvz, p.nz)
        finite = bool(np.isfinite(snap).all())
        # ee0(z, v) = G(v) for this initial state: the z-mean of P_ee is the
        # survival profile sum_z ee / sum_z ee0
        prof = snap.mean(axis=1)
    finally:
        buf.free()
    for d in (d10, d12):
        drifts = {k: _drift(d, d0, k) for k in ("trace", "btrace", "lepton")}
        m0 = math.sqrt(sum((d[f"M0_e{k}"] - d0[f"M0_e{k}"]) ** 2 for k in (1, 2, 3))) / d0["M0_norm"]
        _log(f"HBM-filling run {name}: drift {drifts}, M0 {m0:.3e}, non-finite {d['nonfinite']}")
        assert max(drifts.values()) <= 1e-14
        assert m0 <= 1e-12
        assert d["nonfinite"] == 0
    rp = _rel(prof, pee)
    _log(f"HBM-filling run {name}: snapshot finite {finite}, z-mean vs survival profile {rp:.3e}")
    assert finite and rp <= 1e-12
    assert abs(d10["Pee"] - 1.0) < 1e-3

    # the same physics on a z-subsampled periodic domain against the oracle
    sub = subsampled(workload("c4", mode=C.FAST, scheme=scheme), 8192)
    ps = sub.params
    po = _orc(ps)
    a, b, G1, bG1 = sub.factors()
    s0 = a[:, :, None] * b[:, None, :]
    with _start(ps, a, b, G1, bG1) as sim:
        sim.step(3)
        f3 = sim.download()
    with _start(dataclasses.replace(ps, mode=C.STRICT), a, b, G1, bG1) as sim:
        sim.step(2)
        s2 = sim.download()
    r3 = _rel(f3, O.step(po, s0, 3, moment_form=True))
    eq = np.array_equal(s2, O.step(po, s0, 2))
    _log(f"HBM-filling physics on nz=8192 (dz={ps.dz}): FAST-oracle 3 steps {r3:.3e}; STRICT "
         f"2 steps equal to the oracle: {eq}")
    assert r3 <= 1e-10 and eq


def test_sweep_256_runs(gpu):
    from cosenu_amd.sweep import sweep_runs
    runs, ics = sweep_runs(256)
    nsteps = 20  # one 16-step graph replay plus 4 steps launched one by one
    cells = sum(p.nz * p.nvz for p in runs)
    _log(f"256-run sweep: {cells} cells, nz {min(p.nz for p in runs)}..{max(p.nz for p in runs)}, "
         f"nvz {sorted(set(p.nvz for p in runs))}, {nsteps} steps")

    def ic(r):
        sn, sb, al, sd = ics[r]
        return seeded_ic(runs[r], sn, sb, al, sd)

    with C.Batch(runs) as B:
        info = B.info()
        assert info["nruns"] == 256 and info["cells"] == cells
        for r in range(len(runs)):
            G, bG, s0 = ic(r)
            B.set_profile(r, G[:, :1], bG[:, :1])
            B.upload(r, s0, C.STATE0)
            B.upload(r, s0)
        B.step(nsteps)
        B.sync()
        _log(f"256-run sweep: batch stepped (variant {info['fast_variant']}, "
             f"{info['work_items']} work items of {info['z_chunk']} rows); comparing runs")
        same = 0
        outs = {}
        for r, p in enumerate(runs):
            G, bG, s0 = ic(r)
            out = B.download(r)
            db = B.diagnostics(r)
            with C.Simulation(p) as sim:
                sim.set_profile(G[:, :1], bG[:, :1])
                sim.upload(s0, C.STATE0)
                sim.upload(s0)
                sim.step(nsteps)
                solo = sim.download()
                ds = sim.diagnostics()
            ok = np.array_equal(out, solo) and db == ds
            same += int(ok)
            if not ok:
                _log(f"run {r} (nz={p.nz}, nvz={p.nvz}) differs from its solo run")
            if r in (0, 1):
                outs[r] = (out, s0)
    _log(f"256-run sweep: {same}/256 runs bit-identical to their solo runs")
    assert same == 256
    for r, (out, s0) in outs.items():
        p = runs[r]
        rel = _rel(out, O.step(_orc(p), s0, nsteps, moment_form=True))
        _log(f"256-run sweep: run {r} (nz={p.nz}, nvz={p.nvz}) against the oracle {rel:.3e}")
        assert rel <= 1e-10


def test_surrogate_dataset_grid(gpu):
    """The dataset generator's solver runs (cosenu_amd.surrogate.solve) at the
    surrogate's grid and schedule: a batch of 4 runs, one of them repeated in
    its own context with the same schedule."""
    from cosenu_amd import surrogate as S
    setup = S.SolverSetup(nz=8192, steps=8000, batch=256)
    elns = S.sample_elns(4, seed=77)
    _log(f"surrogate dataset grid: 4 runs, nz={setup.nz}, nvz={S.NV}, {setup.steps} steps")
    Y, secs = S.solve(elns, setup)
    _log(f"surrogate dataset grid: solved in {secs:.1f} s; P_ee(v) range {Y.min():.4f}..{Y.max():.4f}")
    assert Y.shape == (4, S.NV) and np.isfinite(Y).all()
    assert Y.min() > 0.0 and Y.max() < 1.0 + 1e-3
    # run 0 alone: the same time average, bit for bit
    p = S.run_params(setup)
    e = elns[0]
    G, bG, s0 = seeded_ic(p, e.sig_nu, e.sig_nubar, e.alpha, e.seed)
    first = int(round(setup.steps * (1 - setup.avg_frac)))
    acc = np.zeros(S.NV)
    n = 0
    with C.Simulation(p) as sim:
        sim.set_profile(G[:, :1], bG[:, :1])
        sim.upload(s0, C.STATE0)
        sim.upload(s0)
        done = 0
        for target in range(first, setup.steps + 1, setup.sample_every):
            sim.step(target - done)
            done = target
            acc += sim.survival_profile()[0]
            n += 1
    solo = (acc / n).astype(np.float32)
    _log(f"surrogate dataset grid: solo run 0 equal to the batch: {np.array_equal(solo, Y[0])}")
    assert np.array_equal(solo, Y[0])


def test_surrogate_screens_fresh_runs_at_dataset_grid(gpu):
    """The committed ensemble (cosenu_amd/data/surrogate_ensemble.npz) screens
    8 configurations it never saw, at the dataset's grid and schedule: the
    accepted runs take the network's P_ee(v), the rejected ones the solver's;
    the screened profiles are compared with the solver's on every run."""
    import os
    from cosenu_amd import surrogate as S
    path = os.path.join(os.path.dirname(C.__file__), "data", "surrogate_ensemble.npz")
    sur = S.Surrogate.load(path)
    setup = S.SolverSetup(nz=8192, steps=8000)
    elns = S.sample_elns(8, seed=4242)
    X = np.stack([S.features(e) for e in elns])
    pred, accept = sur.screen(X, gate="mean")
    _log(f"surrogate screen: {int(accept.sum())} of 8 accepted (sweep-mean gate)")
    Y, secs = S.solve(elns, setup)
    _log(f"surrogate screen: solver on all 8 runs {secs:.1f} s")
    assert np.isfinite(pred).all() and pred.min() > -0.05 and pred.max() < 1.05
    screened = pred.copy()
    screened[~accept] = Y[~accept]
    err = S.run_errors(screened, Y)
    err_net = S.run_errors(pred, Y)
    _log(f"surrogate screen: per-run |dP_ee| network {np.round(err_net, 4).tolist()}, "
         f"screened mean {err.mean():.4f}")
    assert err[~accept].max(initial=0.0) == 0.0
    # the screened runs' mean error at the level the gate is calibrated for
    # (2% on a sweep's mean; 8 runs leave room for sampling spread)
    assert err.mean() <= 0.04
    for m in sur.members:
        m.close()
