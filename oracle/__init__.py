"""CPU restatement the HIP library is tested against -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package.  The product (``cosenu_amd``) never does.

Two CPU implementations live here:

* ``libcosenu_oracle.so`` -- ``cosenu_oracle.c``, a restatement of the
  reference's RHS / RK4 / initial condition / diagnostics in the reference's
  floating-point operation order.  Always buildable (``make -C oracle``).
* ``_ref/libref_<variant>.so`` -- the reference's own headers
  (src/nuosc.hpp, src/rhs_fv.hpp, src/rhs_fd.hpp, ...) compiled with the
  reference's flags around ``ref_harness.cpp``; one library per preset set.
  Built only where the reference checkout exists (``make -C oracle ref``).

State exchange format everywhere: ``float64[8, nvz, nz]`` with fields
(ee, xx, ex_re, ex_im, bee, bxx, bex_re, bex_im).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libcosenu_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

FV, FD = 0, 1
PERIODIC, OPEN = 0, 1
VAC, MAT, COLL, ADVEC_OFF, KO5 = 1, 2, 4, 8, 16

DIAG_NAMES = ("dP_max", "dP_avg", "dbP_avg", "M0_norm", "M0_e1", "M0_e2", "M0_e3",
              "Pee", "Pbee", "trace", "btrace", "lepton", "nonfinite")


class OrcParams(ctypes.Structure):
    _fields_ = [("scheme", ctypes.c_int), ("bc", ctypes.c_int), ("flags", ctypes.c_int),
                ("nz", ctypes.c_int), ("nvz", ctypes.c_int), ("gz", ctypes.c_int),
                ("z0", ctypes.c_double), ("z1", ctypes.c_double), ("v0", ctypes.c_double),
                ("v1", ctypes.c_double), ("CFL", ctypes.c_double), ("pmo", ctypes.c_double),
                ("omega", ctypes.c_double), ("theta", ctypes.c_double), ("mu", ctypes.c_double),
                ("ko", ctypes.c_double)]


@dataclass
class RunParams:
    """One run's physics and grid, the fields of the reference's job.config
    (parser.hpp:99-256) plus the compile-time presets as runtime flags."""
    nz: int
    nvz: int
    z0: float = -600.0
    z1: float = 600.0
    v0: float = -1.0
    v1: float = 1.0
    CFL: float = 0.4
    scheme: int = FV
    bc: int = PERIODIC
    flags: int = VAC | COLL
    pmo: float = 1.0
    omega: float = 0.0
    theta: float = 1e-6
    mu: float = 1.0
    ko: float = 0.1
    gz: int = -1

    def __post_init__(self):
        if self.gz < 0:
            self.gz = 4 if self.scheme == FV else (3 if self.flags & KO5 else 2)

    def c(self) -> OrcParams:
        return OrcParams(self.scheme, self.bc, self.flags, self.nz, self.nvz, self.gz,
                         self.z0, self.z1, self.v0, self.v1, self.CFL, self.pmo,
                         self.omega, self.theta, self.mu, self.ko)

    def job_config(self, n_iter: int = 20, anal_every: int = 10) -> str:
        """job.config text in the reference's `key : value` syntax."""
        lines = [f"scheme : {'FV' if self.scheme == FV else 'FD'}", f"nz : {self.nz}",
                 f"nvz : {self.nvz}", f"z0 : {self.z0!r}", f"z1 : {self.z1!r}",
                 f"v0 : {self.v0!r}", f"v1 : {self.v1!r}", f"CFL : {self.CFL!r}",
                 f"gz : {self.gz}", f"N_ITER : {n_iter}", f"ANAL_EVERY : {anal_every}",
                 f"pmo : {self.pmo!r}", f"omega : {self.omega!r}", f"theta : {self.theta!r}",
                 f"mu : {self.mu!r}", "perturbation_size : 0.0", "n_dump_rho : 0"]
        return "\n".join(lines) + "\n"


def build(with_ref: bool = True) -> None:
    subprocess.run(["make", "-s", "-C", HERE, "-j8", "all"], check=True)
    if with_ref and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "-j8", "ref"], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.POINTER(OrcParams)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_rhs.argtypes = [P, dp, dp, ctypes.c_int, ctypes.c_int]
        L.orc_step.argtypes = [P, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_init_reference.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, dp, dp, dp]
        L.orc_diagnostics.argtypes = [P, dp, dp, dp, dp, dp]
        L.orc_grid.argtypes = [P, dp, dp, dp]
        L.orc_weno7.argtypes = [dp, dp, dp]
        L.orc_g.argtypes = [ctypes.c_double] * 3
        L.orc_g.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _chk(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def rhs(p: RunParams, state: np.ndarray, moment_form: bool = False, threads: int = 0) -> np.ndarray:
    s = np.ascontiguousarray(state, dtype=np.float64)
    out = np.empty_like(s)
    _chk(lib().orc_rhs(ctypes.byref(p.c()), _dp(s), _dp(out), int(moment_form), threads), "rhs")
    return out


def step(p: RunParams, state: np.ndarray, nsteps: int, moment_form: bool = False,
         threads: int = 0) -> np.ndarray:
    s = np.array(state, dtype=np.float64, order="C", copy=True)
    _chk(lib().orc_step(ctypes.byref(p.c()), _dp(s), nsteps, int(moment_form), threads), "step")
    return s


def init_reference(p: RunParams, perturbation_size: float, signu=0.6, sigbnu=0.6, alpha=0.9):
    G = np.empty((p.nvz, p.nz))
    bG = np.empty((p.nvz, p.nz))
    s = np.empty((8, p.nvz, p.nz))
    _chk(lib().orc_init_reference(ctypes.byref(p.c()), perturbation_size, signu, sigbnu, alpha,
                                  _dp(G), _dp(bG), _dp(s)), "init")
    return G, bG, s


def diagnostics(p: RunParams, G, bG, state0, state) -> dict:
    out = np.zeros(13)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (G, bG, state0, state)]
    _chk(lib().orc_diagnostics(ctypes.byref(p.c()), *[_dp(a) for a in arrs], _dp(out)), "diag")
    return dict(zip(DIAG_NAMES, out.tolist()))


def grid(p: RunParams):
    out = np.zeros(5)
    vz = np.zeros(p.nvz)
    Z = np.zeros(p.nz)
    _chk(lib().orc_grid(ctypes.byref(p.c()), _dp(out), _dp(vz), _dp(Z)), "grid")
    return dict(zip(("dz", "dv", "dt", "ct", "st"), out.tolist())), vz, Z


def weno7(u7) -> tuple[float, float]:
    u = np.ascontiguousarray(u7, dtype=np.float64)
    r = ctypes.c_double()
    l = ctypes.c_double()
    lib().orc_weno7(_dp(u), ctypes.byref(r), ctypes.byref(l))
    return r.value, l.value


# --------------------------------------------------------------------------
# Reference build (oracle/_ref), one library per preset set.

def ref_variant(p: RunParams) -> str:
    """Name of the _ref library whose presets match p (or raise KeyError)."""
    s = "fv" if p.scheme == FV else "fd"
    if p.scheme == FD and p.flags & KO5:
        s += "_ko5"
    if p.bc == OPEN:
        s += "_open"
    f = p.flags & (VAC | MAT | COLL)
    tag = {VAC | COLL: "vc", VAC: "v", COLL: "c", 0: "a", VAC | MAT | COLL: "vmc"}[f]
    s += "_" + tag
    if p.flags & ADVEC_OFF:
        s += "_noadv"
    return s


_ref_cache: dict[str, ctypes.CDLL] = {}


def ref_available(variant: str = "fv_vc") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{variant}.so"))


def ref_lib(variant: str) -> ctypes.CDLL:
    if variant not in _ref_cache:
        path = os.path.join(REF_DIR, f"libref_{variant}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        c = ctypes.c_char_p
        L.ref_rhs.argtypes = [c, dp, dp]
        L.ref_step.argtypes = [c, dp, ctypes.c_int]
        L.ref_initialize.argtypes = [c, dp, dp, dp]
        L.ref_analyse.argtypes = [c, dp, dp, dp, dp, ctypes.c_int, dp]
        L.ref_write_state0.argtypes = [c, dp]
        L.ref_read_state.argtypes = [c, dp]
        L.ref_dump_rho.argtypes = [c, dp, ctypes.c_int]
        _ref_cache[variant] = L
    return _ref_cache[variant]


class RefRun:
    """Runs the reference build for RunParams p inside directory `workdir`
    (the reference reads its config and writes its files relative to it)."""

    def __init__(self, p: RunParams, workdir: str, n_iter: int = 20, anal_every: int = 10):
        self.p = p
        self.dir = workdir
        self.lib = ref_lib(ref_variant(p))
        os.makedirs(workdir, exist_ok=True)
        self.conf = os.path.join(workdir, "job.config")
        with open(self.conf, "w") as f:
            f.write(p.job_config(n_iter, anal_every))

    def _in_dir(self, fn, *args):
        cwd = os.getcwd()
        os.chdir(self.dir)
        try:
            return fn(self.conf.encode(), *args)
        finally:
            os.chdir(cwd)

    def rhs(self, state):
        s = np.ascontiguousarray(state, dtype=np.float64)
        out = np.empty_like(s)
        self._in_dir(self.lib.ref_rhs, _dp(s), _dp(out))
        return out

    def step(self, state, nsteps):
        s = np.array(state, dtype=np.float64, order="C", copy=True)
        self._in_dir(self.lib.ref_step, _dp(s), nsteps)
        return s

    def initialize(self):
        G = np.empty((self.p.nvz, self.p.nz))
        bG = np.empty_like(G)
        s = np.empty((8, self.p.nvz, self.p.nz))
        self._in_dir(self.lib.ref_initialize, _dp(G), _dp(bG), _dp(s))
        return G, bG, s

    def analyse(self, G, bG, state0, state, t):
        m0 = np.zeros(4)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (G, bG, state0, state)]
        self._in_dir(self.lib.ref_analyse, *[_dp(a) for a in arrs], t, _dp(m0))
        return m0

    def write_state0(self, state):
        s = np.ascontiguousarray(state, dtype=np.float64)
        self._in_dir(self.lib.ref_write_state0, _dp(s))

    def read_state(self):
        s = np.empty((8, self.p.nvz, self.p.nz))
        t = self._in_dir(self.lib.ref_read_state, _dp(s))
        return t, s

    def dump_rho(self, state, t):
        s = np.ascontiguousarray(state, dtype=np.float64)
        self._in_dir(self.lib.ref_dump_rho, _dp(s), t)
