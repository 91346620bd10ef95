"""cosenu_amd -- MI355X (gfx950) implementation of COSEnu's time-step hot path.

The compute lives in ``libcosenu_hip.so`` (HIP kernels behind the C ABI of
``include/cosenu_hip.h``).  This module is a thin ctypes binding of that ABI
plus the host-side pieces a Python caller needs: run parameters mirroring the
reference's ``job.config`` (src/parser.hpp:46-97), the initial-condition
generators, and the reference's binary file formats (src/snaps.hpp).

There is no CPU fallback: if the library is missing, importing the
binding's entry points raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field, asdict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcosenu_hip.so")

FV, FD = 0, 1
PERIODIC, OPEN = 0, 1
VAC, MAT, COLL, ADVEC_OFF, KO5 = 1, 2, 4, 8, 16
STRICT, FAST = 0, 1
STATE, STATE0 = 0, 1

FIELDS = ("ee", "xx", "ex_re", "ex_im", "bee", "bxx", "bex_re", "bex_im")


class CosenuError(RuntimeError):
    pass


class _Params(ctypes.Structure):
    _fields_ = [("scheme", ctypes.c_int), ("bc", ctypes.c_int), ("flags", ctypes.c_int),
                ("nz", ctypes.c_int), ("nvz", ctypes.c_int), ("gz", ctypes.c_int),
                ("z0", ctypes.c_double), ("z1", ctypes.c_double), ("v0", ctypes.c_double),
                ("v1", ctypes.c_double), ("CFL", ctypes.c_double), ("pmo", ctypes.c_double),
                ("omega", ctypes.c_double), ("theta", ctypes.c_double), ("mu", ctypes.c_double),
                ("ko", ctypes.c_double), ("mode", ctypes.c_int), ("device", ctypes.c_int)]


class _Info(ctypes.Structure):
    _fields_ = [("dz", ctypes.c_double), ("dv", ctypes.c_double), ("dt", ctypes.c_double),
                ("ct", ctypes.c_double), ("st", ctypes.c_double), ("phy_time", ctypes.c_double),
                ("steps_done", ctypes.c_longlong), ("device_bytes", ctypes.c_size_t),
                ("z_chunk", ctypes.c_int), ("fast_variant", ctypes.c_int),
                ("range_teams", ctypes.c_int)]


class _BatchInfo(ctypes.Structure):
    _fields_ = [("nruns", ctypes.c_int), ("ngroups", ctypes.c_int), ("cells", ctypes.c_longlong),
                ("steps_done", ctypes.c_longlong), ("device_bytes", ctypes.c_size_t),
                ("work_items", ctypes.c_longlong), ("z_chunk", ctypes.c_int),
                ("fast_variant", ctypes.c_int)]


class _Diag(ctypes.Structure):
    _fields_ = [("dP_max", ctypes.c_double), ("dP_avg", ctypes.c_double),
                ("dbP_avg", ctypes.c_double), ("M0_norm", ctypes.c_double),
                ("M0", ctypes.c_double * 3), ("Pee", ctypes.c_double), ("Pbee", ctypes.c_double),
                ("trace", ctypes.c_double), ("btrace", ctypes.c_double),
                ("lepton", ctypes.c_double), ("nonfinite", ctypes.c_double)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_planes_t = _dp * 8


def lib() -> ctypes.CDLL:
    """Load libcosenu_hip.so (raises CosenuError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CosenuError(f"{LIB_PATH} is missing; build it with `make -C cosenu_amd/csrc` "
                          "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.cosenu_last_error.restype = ctypes.c_char_p
    L.cosenu_create.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(vp)]
    L.cosenu_destroy.argtypes = [vp]
    L.cosenu_get_info.argtypes = [vp, ctypes.POINTER(_Info)]
    L.cosenu_set_profile.argtypes = [vp, _dp, _dp, ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_upload_planes.argtypes = [vp, ctypes.c_int, _planes_t, ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_download_planes.argtypes = [vp, ctypes.c_int, _planes_t, ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_upload_interleaved.argtypes = [vp, ctypes.c_int, _dp]
    L.cosenu_download_interleaved.argtypes = [vp, ctypes.c_int, _dp]
    L.cosenu_step.argtypes = [vp, ctypes.c_int]
    L.cosenu_prepare.argtypes = [vp]
    L.cosenu_sync.argtypes = [vp]
    L.cosenu_rhs.argtypes = [vp, _planes_t, ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_diagnostics.argtypes = [vp, ctypes.POINTER(_Diag)]
    L.cosenu_survival_profile.argtypes = [vp, _dp, _dp]
    L.cosenu_snapshot_pee_async.argtypes = [vp, vp]
    L.cosenu_snapshot_wait.argtypes = [vp]
    L.cosenu_host_alloc.argtypes = [ctypes.c_size_t]
    L.cosenu_host_alloc.restype = vp
    L.cosenu_host_free.argtypes = [vp]
    L.cosenu_set_timing.argtypes = [vp, ctypes.c_int]
    L.cosenu_get_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_longlong)]
    L.cosenu_stage_kernel_name.argtypes = [vp, ctypes.c_int]
    L.cosenu_stage_kernel_name.restype = ctypes.c_char_p
    L.cosenu_copy_state.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cosenu_init_separable.argtypes = [vp, ctypes.c_int, _dp, _dp]
    L.cosenu_set_profile_v.argtypes = [vp, _dp, _dp]
    ip = ctypes.POINTER(ctypes.c_int)
    L.cosenu_autotune.argtypes = [ctypes.POINTER(_Params), ip, ip, ctypes.POINTER(ctypes.c_double)]
    L.cosenu_batch_create.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, ctypes.POINTER(vp)]
    L.cosenu_batch_destroy.argtypes = [vp]
    L.cosenu_batch_get_info.argtypes = [vp, ctypes.POINTER(_BatchInfo)]
    L.cosenu_batch_autotune.argtypes = [ctypes.POINTER(_Params), ctypes.c_int, ip,
                                        ctypes.POINTER(ctypes.c_double)]
    L.cosenu_batch_set_profile.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_batch_upload_planes.argtypes = [vp, ctypes.c_int, ctypes.c_int, _planes_t,
                                             ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_batch_download_planes.argtypes = [vp, ctypes.c_int, ctypes.c_int, _planes_t,
                                               ctypes.c_ssize_t, ctypes.c_int]
    L.cosenu_batch_step.argtypes = [vp, ctypes.c_int]
    L.cosenu_batch_sync.argtypes = [vp]
    L.cosenu_batch_diagnostics.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Diag)]
    L.cosenu_batch_survival_profile.argtypes = [vp, ctypes.c_int, _dp, _dp]
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().cosenu_last_error().decode(errors="replace")
        raise CosenuError(f"{what} failed ({rc}): {msg}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _need(ok: bool, what: str):
    """Shape checks before raw pointers reach the library (kept under -O)."""
    if not ok:
        raise ValueError(what)


@dataclass
class RunParams:
    """A run's grid and physics: the keys of the reference's job.config
    (parser.hpp:99-256) plus its compile-time presets as runtime fields."""
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
    mode: int = FAST
    device: int = 0

    def __post_init__(self):
        if self.gz < 0:
            self.gz = 4 if self.scheme == FV else (3 if self.flags & KO5 else 2)

    def c(self) -> _Params:
        return _Params(self.scheme, self.bc, self.flags, self.nz, self.nvz, self.gz, self.z0,
                       self.z1, self.v0, self.v1, self.CFL, self.pmo, self.omega, self.theta,
                       self.mu, self.ko, self.mode, self.device)

    @property
    def dz(self) -> float:
        return (self.z1 - self.z0) / self.nz

    @property
    def dv(self) -> float:
        return (self.v1 - self.v0) / self.nvz

    @property
    def dt(self) -> float:
        return self.CFL * self.dz

    def vz(self) -> np.ndarray:
        return self.v0 + (np.arange(self.nvz) + 0.5) * self.dv

    def Z(self) -> np.ndarray:
        return self.z0 + (np.arange(self.nz) + 0.5) * self.dz


class Simulation:
    """One state on one GPU: the role of the reference's NuOsc
    (src/nuosc.hpp:7-191).  States are exchanged as float64[8, nvz, nz]."""

    def __init__(self, p: RunParams):
        self.p = p
        h = ctypes.c_void_p()
        _check(lib().cosenu_create(ctypes.byref(p.c()), ctypes.byref(h)), "cosenu_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cosenu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- data in/out ---------------------------------------------------------
    def _planes(self, arr: np.ndarray):
        _need(arr.shape == (8, self.p.nvz, self.p.nz) and arr.dtype == np.float64
              and arr.flags.c_contiguous, "state must be C-contiguous float64 [8, nvz, nz]")
        return _planes_t(*[arr[f].ctypes.data_as(_dp) for f in range(8)])

    def set_profile(self, G: np.ndarray, bG: np.ndarray):
        G = np.ascontiguousarray(np.broadcast_to(G, (self.p.nvz, self.p.nz)), dtype=np.float64)
        bG = np.ascontiguousarray(np.broadcast_to(bG, (self.p.nvz, self.p.nz)), dtype=np.float64)
        _check(lib().cosenu_set_profile(self._h, _ptr(G), _ptr(bG), self.p.nz, 0),
               "cosenu_set_profile")

    def upload(self, state: np.ndarray, which: int = STATE):
        s = np.ascontiguousarray(state, dtype=np.float64)
        _check(lib().cosenu_upload_planes(self._h, which, self._planes(s), self.p.nz, 0),
               "cosenu_upload_planes")

    def download(self, which: int = STATE) -> np.ndarray:
        out = np.empty((8, self.p.nvz, self.p.nz))
        _check(lib().cosenu_download_planes(self._h, which, self._planes(out), self.p.nz, 0),
               "cosenu_download_planes")
        return out

    def upload_interleaved(self, buf: np.ndarray, which: int = STATE):
        b = np.ascontiguousarray(buf, dtype=np.float64)
        _need(b.size == 8 * self.p.nvz * self.p.nz, "interleaved state must hold 8*nvz*nz values")
        _check(lib().cosenu_upload_interleaved(self._h, which, _ptr(b)), "upload_interleaved")

    def download_interleaved(self, which: int = STATE) -> np.ndarray:
        b = np.empty((self.p.nvz, self.p.nz, 8))
        _check(lib().cosenu_download_interleaved(self._h, which, _ptr(b)), "download_interleaved")
        return b

    # -- compute ---------------------------------------------------------------
    def step(self, nsteps: int = 1):
        _check(lib().cosenu_step(self._h, int(nsteps)), "cosenu_step")

    def prepare(self):
        """Capture the step graph now, so the next long step() pays no capture."""
        _check(lib().cosenu_prepare(self._h), "cosenu_prepare")

    def sync(self):
        _check(lib().cosenu_sync(self._h), "cosenu_sync")

    def rhs(self) -> np.ndarray:
        out = np.empty((8, self.p.nvz, self.p.nz))
        _check(lib().cosenu_rhs(self._h, self._planes(out), self.p.nz, 0), "cosenu_rhs")
        return out

    def diagnostics(self) -> dict:
        d = _Diag()
        _check(lib().cosenu_diagnostics(self._h, ctypes.byref(d)), "cosenu_diagnostics")
        return _diag_dict(d)

    def survival_profile(self):
        a = np.empty(self.p.nvz)
        b = np.empty(self.p.nvz)
        _check(lib().cosenu_survival_profile(self._h, _ptr(a), _ptr(b)), "survival_profile")
        return a, b

    def info(self) -> dict:
        i = _Info()
        _check(lib().cosenu_get_info(self._h, ctypes.byref(i)), "cosenu_get_info")
        return {k: getattr(i, k) for k, _ in _Info._fields_}

    def init_separable(self, a: np.ndarray, b: np.ndarray, which: int = STATE):
        """state[f, v, z] = a[f, v] * b[f, z] built on the device."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        _need(a.shape == (8, self.p.nvz) and b.shape == (8, self.p.nz),
              "separable factors must be [8, nvz] and [8, nz]")
        _check(lib().cosenu_init_separable(self._h, which, _ptr(a), _ptr(b)), "init_separable")

    def set_profile_v(self, G: np.ndarray, bG: np.ndarray):
        G = np.ascontiguousarray(G, dtype=np.float64)
        bG = np.ascontiguousarray(bG, dtype=np.float64)
        _need(G.shape == (self.p.nvz,) and bG.shape == (self.p.nvz,),
              "profiles must have nvz values each")
        _check(lib().cosenu_set_profile_v(self._h, _ptr(G), _ptr(bG)), "set_profile_v")

    def copy_state(self, src: int, dst: int):
        _check(lib().cosenu_copy_state(self._h, src, dst), "cosenu_copy_state")

    def snapshot_pee_async(self, buf: "PinnedBuffer"):
        _need(buf.n >= self.p.nvz * self.p.nz, "snapshot buffer must hold nvz*nz values")
        _check(lib().cosenu_snapshot_pee_async(self._h, buf.ptr), "snapshot_pee_async")

    def snapshot_wait(self):
        _check(lib().cosenu_snapshot_wait(self._h), "snapshot_wait")

    def set_timing(self, on: bool):
        _check(lib().cosenu_set_timing(self._h, int(on)), "cosenu_set_timing")

    def timing(self):
        ms = (ctypes.c_double * 4)()
        cnt = (ctypes.c_longlong * 4)()
        _check(lib().cosenu_get_timing(self._h, ms, cnt), "cosenu_get_timing")
        return list(ms), list(cnt)

    def stage_kernel_name(self, stage: int) -> str:
        return lib().cosenu_stage_kernel_name(self._h, stage).decode()


def _diag_dict(d: _Diag) -> dict:
    return {"dP_max": d.dP_max, "dP_avg": d.dP_avg, "dbP_avg": d.dbP_avg,
            "M0_norm": d.M0_norm, "M0_e1": d.M0[0], "M0_e2": d.M0[1], "M0_e3": d.M0[2],
            "Pee": d.Pee, "Pbee": d.Pbee, "trace": d.trace, "btrace": d.btrace,
            "lepton": d.lepton, "nonfinite": d.nonfinite}


def autotune(p: RunParams):
    """Time every FAST kernel variant and z chunk for p's shape; the winner is
    used by contexts created afterwards.  Returns (variant, z_chunk, ms)."""
    v, L, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    _check(lib().cosenu_autotune(ctypes.byref(p.c()), ctypes.byref(v), ctypes.byref(L),
                                 ctypes.byref(ms)), "cosenu_autotune")
    return v.value, L.value, ms.value


def batch_autotune(runs: list):
    """Time the work-item chunk lengths of a batch of these runs (on a
    scratch batch); the winner is used by batches created afterwards.
    Returns (z_chunk, ms per batch step)."""
    arr = (_Params * len(runs))(*[r.c() for r in runs])
    L, ms = ctypes.c_int(), ctypes.c_double()
    _check(lib().cosenu_batch_autotune(arr, len(runs), ctypes.byref(L), ctypes.byref(ms)),
           "cosenu_batch_autotune")
    return L.value, ms.value


class Batch:
    """Many independent FAST-mode runs advanced together (cosenu_batch_*)."""

    def __init__(self, runs: list):
        self.runs = list(runs)
        arr = (_Params * len(self.runs))(*[r.c() for r in self.runs])
        h = ctypes.c_void_p()
        _check(lib().cosenu_batch_create(arr, len(self.runs), ctypes.byref(h)), "batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cosenu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _planes(self, r: int, arr: np.ndarray):
        p = self.runs[r]
        _need(arr.shape == (8, p.nvz, p.nz) and arr.dtype == np.float64 and arr.flags.c_contiguous,
              "state must be C-contiguous float64 [8, nvz, nz]")
        return _planes_t(*[arr[f].ctypes.data_as(_dp) for f in range(8)])

    def set_profile(self, r: int, G, bG):
        p = self.runs[r]
        G = np.ascontiguousarray(np.broadcast_to(G, (p.nvz, p.nz)), dtype=np.float64)
        bG = np.ascontiguousarray(np.broadcast_to(bG, (p.nvz, p.nz)), dtype=np.float64)
        _check(lib().cosenu_batch_set_profile(self._h, r, _ptr(G), _ptr(bG), p.nz, 0),
               "batch_set_profile")

    def upload(self, r: int, state: np.ndarray, which: int = STATE):
        s = np.ascontiguousarray(state, dtype=np.float64)
        _check(lib().cosenu_batch_upload_planes(self._h, r, which, self._planes(r, s),
                                                self.runs[r].nz, 0), "batch_upload")

    def download(self, r: int, which: int = STATE) -> np.ndarray:
        p = self.runs[r]
        out = np.empty((8, p.nvz, p.nz))
        _check(lib().cosenu_batch_download_planes(self._h, r, which, self._planes(r, out), p.nz, 0),
               "batch_download")
        return out

    def step(self, nsteps: int = 1):
        _check(lib().cosenu_batch_step(self._h, int(nsteps)), "batch_step")

    def sync(self):
        _check(lib().cosenu_batch_sync(self._h), "batch_sync")

    def diagnostics(self, r: int) -> dict:
        d = _Diag()
        _check(lib().cosenu_batch_diagnostics(self._h, r, ctypes.byref(d)), "batch_diagnostics")
        return _diag_dict(d)

    def survival_profile(self, r: int):
        p = self.runs[r]
        a, b = np.empty(p.nvz), np.empty(p.nvz)
        _check(lib().cosenu_batch_survival_profile(self._h, r, _ptr(a), _ptr(b)), "batch_survival")
        return a, b

    def info(self) -> dict:
        i = _BatchInfo()
        _check(lib().cosenu_batch_get_info(self._h, ctypes.byref(i)), "batch_get_info")
        return {k: getattr(i, k) for k, _ in _BatchInfo._fields_}


class PinnedBuffer:
    """Page-locked host memory from cosenu_host_alloc, viewed as float64."""

    def __init__(self, n: int):
        self.n = n
        self.ptr = lib().cosenu_host_alloc(n * 8)
        if not self.ptr:
            raise CosenuError("cosenu_host_alloc failed")
        self.array = np.ctypeslib.as_array(ctypes.cast(self.ptr, _dp), shape=(n,))

    def free(self):
        if self.ptr:
            lib().cosenu_host_free(self.ptr)
            self.ptr = None
