"""Learned surrogate for electron-lepton-number (ELN) sweeps.

Maps an initial ELN angular distribution to the asymptotic, angle-resolved
survival probability P_ee(v) that the FP64 solver produces, so that a sweep
over many ELN configurations can skip most solver runs.

* Features (67): g_nu(v) - g_nubar(v) at the 64 velocity-bin centres of
  [-1, 1] (g_nu = g(v; 1, sigma_nu), g_nubar = alpha g(v; 1, sigma_nubar), the
  normalised Gaussian profile of the reference's initial condition,
  src/miscellaneous_funcs.hpp:57-62), then alpha, sigma_nu, sigma_nubar.
* Targets (64): P_ee(v) = sum_z ee(z, v) / sum_z ee0(z, v), averaged over the
  final 20% of the run's steps, from batched FAST-mode solver runs
  (cosenu_batch_*).  Only solver output is used for training.
* Network: 67 -> 256 -> 256 -> 256 -> 64, GELU, BF16 weights and activations
  with FP32 accumulation on the matrix cores; forward, backward and AdamW are
  the hand-written kernels of csrc/mlp.hip (cosenu_mlp_*).
* Fallback: an ensemble of independently initialised networks; the mean
  over bins of the ensemble standard deviation is the per-run error
  indicator.  Two acceptance thresholds are calibrated on solver runs held
  out of training and measured on a second held-out set: a per-run gate (no
  calibration run whose own error, the mean over bins of |dP_ee|, exceeds
  the bound (2%) is accepted) and a sweep-mean gate (the screened sweep's
  mean error stays within the bound).  Every rejected run is run by the
  solver.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import (COLL, FAST, FV, VAC, Batch, CosenuError, RunParams, lib)
from .initial import g_profile, seeded_ic

DIMS = (67, 256, 256, 256, 64)
NV = 64


# ---------------------------------------------------------------------------
# ctypes binding of cosenu_mlp_*

_fp = ctypes.POINTER(ctypes.c_float)


def _mlib():
    L = lib()
    if not getattr(L, "_mlp_bound", False):
        vp = ctypes.c_void_p
        L.cosenu_mlp_last_error.restype = ctypes.c_char_p
        L.cosenu_mlp_create.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.POINTER(vp)]
        L.cosenu_mlp_destroy.argtypes = [vp]
        L.cosenu_mlp_num_params.argtypes = [vp]
        L.cosenu_mlp_num_params.restype = ctypes.c_longlong
        L.cosenu_mlp_set_params.argtypes = [vp, _fp]
        L.cosenu_mlp_get_params.argtypes = [vp, _fp]
        L.cosenu_mlp_forward.argtypes = [vp, _fp, ctypes.c_int, _fp]
        L.cosenu_mlp_gradients.argtypes = [vp, _fp, _fp, ctypes.c_int, _fp, _fp]
        L.cosenu_mlp_train_step.argtypes = [vp, _fp, _fp, ctypes.c_int] + [ctypes.c_float] * 5 + [_fp]
        L.cosenu_mlp_sync.argtypes = [vp]
        L._mlp_bound = True
    return L


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class MLP:
    """The network on the GPU (cosenu_mlp_*)."""

    def __init__(self, dims=DIMS, max_batch: int = 4096, device: int = 0):
        self.dims = tuple(int(d) for d in dims)
        L = _mlib()
        arr = (ctypes.c_int * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        self._chk(L.cosenu_mlp_create(arr, len(self.dims) - 1, max_batch, device, ctypes.byref(h)),
                  "create")
        self._h = h
        self.max_batch = max_batch
        self.nparams = L.cosenu_mlp_num_params(h)

    def _chk(self, rc, what):
        if rc != 0:
            raise CosenuError(f"mlp {what} failed ({rc}): "
                              f"{_mlib().cosenu_mlp_last_error().decode(errors='replace')}")

    def close(self):
        if getattr(self, "_h", None):
            _mlib().cosenu_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, flat):
        flat = _f32(flat)
        if flat.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters, got {flat.size}")
        self._chk(_mlib().cosenu_mlp_set_params(self._h, _p(flat)), "set_params")

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        self._chk(_mlib().cosenu_mlp_get_params(self._h, _p(out)), "get_params")
        return out

    def forward(self, X) -> np.ndarray:
        X = _f32(X)
        out = np.empty((X.shape[0], self.dims[-1]), dtype=np.float32)
        for s in range(0, X.shape[0], self.max_batch):
            xs = np.ascontiguousarray(X[s:s + self.max_batch])
            ys = np.empty((xs.shape[0], self.dims[-1]), dtype=np.float32)
            self._chk(_mlib().cosenu_mlp_forward(self._h, _p(xs), xs.shape[0], _p(ys)), "forward")
            out[s:s + xs.shape[0]] = ys
        return out

    def gradients(self, X, T):
        X, T = _f32(X), _f32(T)
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        self._chk(_mlib().cosenu_mlp_gradients(self._h, _p(X), _p(T), X.shape[0], _p(g),
                                               ctypes.byref(loss)), "gradients")
        return float(loss.value), g

    def train_step(self, X, T, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0,
                   want_loss=False):
        X, T = _f32(X), _f32(T)
        loss = ctypes.c_float()
        self._chk(_mlib().cosenu_mlp_train_step(self._h, _p(X), _p(T), X.shape[0], lr, beta1, beta2,
                                                eps, weight_decay,
                                                ctypes.byref(loss) if want_loss else None),
                  "train_step")
        return float(loss.value) if want_loss else None


def init_params(dims=DIMS, seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, flat FP32 (layout of cosenu_mlp_set_params)."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, n in zip(dims[:-1], dims[1:]):
        lim = math.sqrt(6.0 / k)
        parts.append(rng.uniform(-lim, lim, k * n))
        parts.append(np.zeros(n))
    return np.concatenate(parts).astype(np.float32)


# ---------------------------------------------------------------------------
# Dataset

@dataclass
class ELN:
    sig_nu: float
    sig_nubar: float
    alpha: float
    seed: int


@dataclass
class SolverSetup:
    """Grid and schedule of the solver runs behind the data."""
    nz: int = 2048
    z0: float = -600.0
    z1: float = 600.0
    steps: int = 2000
    avg_frac: float = 0.2   # average P_ee(v) over the final 20% of the steps
    sample_every: int = 20
    batch: int = 256


def features(e: ELN, nv: int = NV) -> np.ndarray:
    dv = 2.0 / nv
    v = -1.0 + (np.arange(nv) + 0.5) * dv
    gn = np.array([g_profile(float(x), 1.0, e.sig_nu) for x in v])
    gb = np.array([e.alpha * g_profile(float(x), 1.0, e.sig_nubar) for x in v])
    return np.concatenate([gn - gb, [e.alpha, e.sig_nu, e.sig_nubar]]).astype(np.float32)


def sample_elns(n: int, seed: int) -> list:
    rng = np.random.default_rng(seed)
    return [ELN(float(rng.uniform(0.4, 0.8)), float(rng.uniform(0.4, 0.8)),
                float(rng.uniform(0.7, 1.1)), int(seed * 1000003 + i)) for i in range(n)]


def run_params(setup: SolverSetup) -> RunParams:
    return RunParams(nz=setup.nz, nvz=NV, z0=setup.z0, z1=setup.z1, CFL=0.4, scheme=FV,
                     flags=VAC | COLL, omega=0.0, theta=1e-6, pmo=1.0, mu=1.0, mode=FAST)


def solve(elns: list, setup: SolverSetup, log=None) -> tuple:
    """FP64 solver targets for each ELN: the time-averaged P_ee(v).
    Returns (targets float32[n, 64], seconds of GPU solver time)."""
    p = run_params(setup)
    out = np.empty((len(elns), NV), dtype=np.float32)
    t_total = 0.0
    first = int(round(setup.steps * (1 - setup.avg_frac)))
    for b0 in range(0, len(elns), setup.batch):
        chunk = elns[b0:b0 + setup.batch]
        with Batch([p] * len(chunk)) as B:
            for r, e in enumerate(chunk):
                G, bG, s0 = seeded_ic(p, e.sig_nu, e.sig_nubar, e.alpha, e.seed)
                B.set_profile(r, G[:, :1], bG[:, :1])
                B.upload(r, s0, 1)
                B.upload(r, s0, 0)
            B.sync()
            t0 = time.perf_counter()
            acc = np.zeros((len(chunk), NV))
            nsamp = 0
            done = 0
            while log and first - done > 1000:  # progress lines on long runs
                B.step(1000)
                done += 1000
                B.sync()
                log(f"solver: runs {b0}..{b0 + len(chunk)}: step {done}")
            for target in range(first, setup.steps + 1, setup.sample_every):
                B.step(target - done)
                done = target
                for r in range(len(chunk)):
                    acc[r] += B.survival_profile(r)[0]
                nsamp += 1
            B.sync()
            t_total += time.perf_counter() - t0
            out[b0:b0 + len(chunk)] = (acc / nsamp).astype(np.float32)
        if log:
            log(f"solver: {b0 + len(chunk)}/{len(elns)} runs, {t_total:.1f} s")
    return out, t_total


# ---------------------------------------------------------------------------
# Training, ensemble, fallback

@dataclass
class TrainConfig:
    steps: int = 3000
    batch: int = 256
    lr: float = 2e-3
    lr_min: float = 2e-5
    weight_decay: float = 1e-4
    members: int = 4


def train_member(X, Y, cfg: TrainConfig, seed: int, net: MLP | None = None, log=None):
    net = net or MLP(max_batch=max(cfg.batch, 4096))
    net.set_params(init_params(DIMS, seed))
    rng = np.random.default_rng(seed + 7)
    n = X.shape[0]
    losses = []
    for s in range(cfg.steps):
        idx = rng.choice(n, size=min(cfg.batch, n), replace=False)
        lr = cfg.lr_min + 0.5 * (cfg.lr - cfg.lr_min) * (1 + math.cos(math.pi * s / cfg.steps))
        want = (s % 500 == 0) or s == cfg.steps - 1
        l = net.train_step(X[idx], Y[idx], lr, weight_decay=cfg.weight_decay, want_loss=want)
        if want:
            losses.append((s, l))
            if log:
                log(f"member {seed}: step {s} loss {l:.3e}")
    return net, losses


def run_errors(pred, Y):
    """Per-run error: the mean over the velocity bins of |dP_ee|."""
    return np.abs(np.asarray(pred, dtype=np.float64) - Y).mean(axis=1)


class Surrogate:
    """Ensemble of trained networks with two acceptance gates.

    Both gate on the ensemble spread (mean over bins of the members'
    standard deviation) with thresholds set by `calibrate` on solver runs the
    networks were not trained on; every rejected run goes to the solver.

    * per-run gate (tau): just below the smallest spread of a calibration
      run whose own error (mean over bins of |dP_ee|) exceeds the bound, so
      that no calibration run above the bound is accepted;
    * sweep-mean gate (tau_mean): the largest threshold at which the
      screened calibration sweep (accepted runs at their network error,
      rejected runs at the solver's, i.e. zero) keeps its mean error within
      the bound.

    `evaluate` measures both on further unseen runs (accept rate, largest
    accepted error, recall of the runs above the bound, screened mean)."""

    def __init__(self, members: list, bound: float = 0.02, meta: dict | None = None):
        self.members = members
        self.bound = bound
        self.tau = -1.0       # per-run gate threshold on the ensemble spread
        self.tau_mean = -1.0  # sweep-mean gate threshold
        self.meta = dict(meta or {})

    def predict(self, X):
        P = np.stack([m.forward(X) for m in self.members])
        return P.mean(axis=0), P.std(axis=0).mean(axis=1)

    def calibrate(self, X, Y):
        mean, spread = self.predict(X)
        err = run_errors(mean, Y)
        bad = err > self.bound
        if bad.any():
            self.tau = float(np.nextafter(spread[bad].min(), -np.inf))
        else:
            self.tau = float(spread.max())
        order = np.argsort(spread, kind="stable")
        screened_mean = np.cumsum(err[order]) / len(err)  # accept the k least spread
        ok = np.nonzero(screened_mean <= self.bound)[0]
        self.tau_mean = float(spread[order][ok[-1]]) if ok.size else -1.0
        out = {"runs": int(len(err)), "above_bound": int(bad.sum()),
               "mean_err_all": float(err.mean()), "tau": self.tau, "tau_mean": self.tau_mean}
        for g, t in (("run", self.tau), ("mean", self.tau_mean)):
            acc = spread <= t
            out[f"{g}_gate_accept_rate"] = float(acc.mean())
            out[f"{g}_gate_max_accepted_run_err"] = float(err[acc].max()) if acc.any() else None
        return out

    def evaluate(self, X, Y):
        """Both gates on unseen runs: rejected runs take the solver's values."""
        mean, spread = self.predict(X)
        err = run_errors(mean, Y)
        bad = err > self.bound
        out = {"runs": int(len(err)), "above_bound": int(bad.sum()),
               "mean_err_network_only": float(err.mean())}
        for g, t in (("run", self.tau), ("mean", self.tau_mean)):
            acc = spread <= t
            out[f"{g}_gate"] = {
                "accept_rate": float(acc.mean()),
                "max_accepted_run_err": float(err[acc].max()) if acc.any() else None,
                "mean_accepted_run_err": float(err[acc].mean()) if acc.any() else None,
                "recall_of_above_bound": float((bad & ~acc).sum() / bad.sum()) if bad.any() else None,
                "accepted_above_bound": int((bad & acc).sum()),
                "mean_err_screened": float(np.where(acc, err, 0.0).mean())}
        return out

    def screen(self, X, gate: str = "run"):
        """Network prediction and the accept mask of a gate ("run" or "mean")."""
        mean, spread = self.predict(X)
        return mean, spread <= (self.tau if gate == "run" else self.tau_mean)

    def save(self, path: str):
        """Ensemble, threshold and metadata in one .npz (no pickles)."""
        import json
        params = np.stack([m.get_params() for m in self.members])
        np.savez_compressed(path, params=params, dims=np.array(self.members[0].dims),
                            tau=np.float64(self.tau), tau_mean=np.float64(self.tau_mean),
                            bound=np.float64(self.bound), meta=json.dumps(self.meta))

    @classmethod
    def load(cls, path: str, max_batch: int = 4096, device: int = 0) -> "Surrogate":
        import json
        with np.load(path, allow_pickle=False) as z:
            dims = tuple(int(d) for d in z["dims"])
            members = []
            for flat in z["params"]:
                m = MLP(dims, max_batch=max_batch, device=device)
                m.set_params(flat)
                members.append(m)
            s = cls(members, bound=float(z["bound"]), meta=json.loads(str(z["meta"])))
            s.tau = float(z["tau"])
            s.tau_mean = float(z["tau_mean"])
        return s
