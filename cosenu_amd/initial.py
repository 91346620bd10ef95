"""Initial conditions and the named workloads.

``reference_ic`` is the reference's NuOsc::initialize (src/initialize.hpp:6-47
with g/eps/eps_ of src/miscellaneous_funcs.hpp:57-77), evaluated with the C
library's exp/erf/sqrt per velocity bin and per z cell (the same libm calls
the reference makes), so it reproduces the reference's initial state bit for
bit.  ``seeded_ic`` builds the flavour-unstable states of the benchmark
workloads: G = g(v; 1, sigma_nu), bG = alpha g(v; 1, sigma_nubar),
ee = G, bee = bG, xx = bxx = 0, ex = amp G e^{i phi_j} (+ an optional Gaussian
seed amp exp(-z^2/50) G), the same phase field phi_j for nu and nubar, drawn
from SplitMix64 with a recorded seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from . import RunParams, FV, FD, VAC, COLL, FAST, OPEN


def _vz(p) -> np.ndarray:
    dv = (p.v1 - p.v0) / (p.nvz)
    return p.v0 + (np.arange(p.nvz) + 0.5) * dv


def _Z(p) -> np.ndarray:
    dz = (p.z1 - p.z0) / p.nz
    return p.z0 + (np.arange(p.nz) + 0.5) * dz


def g_profile(v: float, v0: float, sigma: float) -> float:
    """Normalised Gaussian angular profile on [-1, 1] (miscellaneous_funcs.hpp:57-62)."""
    ex = (v - v0) * (v - v0) / (2.0 * sigma * sigma)
    N = sigma * math.sqrt(math.pi / 2.0) * (math.erf((1.0 + v0) / sigma / math.sqrt(2.0))
                                            + math.erf((1.0 - v0) / sigma / math.sqrt(2.0)))
    return math.exp(-ex) / N


def reference_ic(p: RunParams, perturbation_size: float, signu: float = 0.6,
                 sigbnu: float = 0.6, alpha: float = 0.9):
    """(G, bG, state) of NuOsc::initialize; G, bG are [nvz, nz], state [8, nvz, nz]."""
    vz = _vz(p)
    Z = _Z(p)
    G1 = np.array([g_profile(float(v), 1.0, signu) for v in vz])
    bG1 = np.array([alpha * g_profile(float(v), 1.0, sigbnu) for v in vz])
    e = np.array([perturbation_size * math.exp(-(float(z) - 0.0) * (float(z) - 0.0) / 50.0)
                  for z in Z])
    e_ = np.array([math.sqrt(1.0 - x * x) for x in e])
    G = np.repeat(G1[:, None], p.nz, axis=1)
    bG = np.repeat(bG1[:, None], p.nz, axis=1)
    s = np.empty((8, p.nvz, p.nz))
    s[0] = 0.5 * G * (1.0 + e_)[None, :]
    s[1] = 0.5 * G * (1.0 - e_)[None, :]
    s[2] = 0.5 * G * (0.0 + e)[None, :]
    s[3] = -0.0
    s[4] = 0.5 * bG * (1.0 + e_)[None, :]
    s[5] = 0.5 * bG * (1.0 - e_)[None, :]
    s[6] = 0.5 * bG * (0.0 + e)[None, :]
    s[7] = 0.0
    return G, bG, s


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from SplitMix64 (Steele, Lea, Flood 2014): the top 53
    bits of each output times 2^-53."""
    out = np.empty(n, dtype=np.uint64)
    x = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
    g = np.uint64(0x9E3779B97F4A7C15)
    m1 = np.uint64(0xBF58476D1CE4E5B9)
    m2 = np.uint64(0x94D049BB133111EB)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + idx * g
        z = (z ^ (z >> np.uint64(30))) * m1
        z = (z ^ (z >> np.uint64(27))) * m2
        z = z ^ (z >> np.uint64(31))
    out[:] = z
    return (out >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def seeded_factors(p: RunParams, sig_nu: float, sig_nubar: float, alpha: float, seed: int,
                   amp: float = 1e-6, gaussian_seed: bool = False):
    """Separable form of the seeded state: field f at (v, z) is a[f, v] * b[f, z].
    Returns (a [8, nvz], b [8, nz], G(v), bG(v))."""
    vz = _vz(p)
    G1 = np.array([g_profile(float(v), 1.0, sig_nu) for v in vz])
    bG1 = np.array([alpha * g_profile(float(v), 1.0, sig_nubar) for v in vz])
    # libm per cell (not numpy's SIMD kernels, whose last bit can depend on
    # the host CPU): the state is then identical on every x86-64 host.
    phi = (2.0 * math.pi * splitmix64_uniform(seed, p.nz)).tolist()
    c = np.array([math.cos(x) for x in phi])
    s_ = np.array([math.sin(x) for x in phi])
    if gaussian_seed:
        gs = np.array([math.exp(-(z * z) / 50.0) for z in _Z(p).tolist()])
        c = c + gs
    one = np.ones(p.nz)
    zero = np.zeros(p.nz)
    a = np.stack([G1, np.zeros_like(G1), amp * G1, amp * G1,
                  bG1, np.zeros_like(G1), amp * bG1, amp * bG1])
    b = np.stack([one, zero, c, s_, one, zero, c, s_])
    return a, b, G1, bG1


def seeded_ic(p: RunParams, sig_nu: float, sig_nubar: float, alpha: float, seed: int,
              amp: float = 1e-6, gaussian_seed: bool = False):
    """(G, bG, state) of a flavour-unstable ELN state with random phases per z."""
    a, b, G1, bG1 = seeded_factors(p, sig_nu, sig_nubar, alpha, seed, amp, gaussian_seed)
    st = a[:, :, None] * b[:, None, :]
    G = np.repeat(G1[:, None], p.nz, axis=1)
    bG = np.repeat(bG1[:, None], p.nz, axis=1)
    return G, bG, st


@dataclass
class Workload:
    name: str
    params: RunParams
    steps: int
    sig_nu: float
    sig_nubar: float
    alpha: float
    seed: int = 20261203
    gaussian_seed: bool = False

    def ic(self):
        return seeded_ic(self.params, self.sig_nu, self.sig_nubar, self.alpha, self.seed,
                         gaussian_seed=self.gaussian_seed)

    def factors(self):
        return seeded_factors(self.params, self.sig_nu, self.sig_nubar, self.alpha, self.seed,
                              gaussian_seed=self.gaussian_seed)


def workload(name: str, mode: int = FAST, scheme: int = FV) -> Workload:
    """The named benchmark workloads.  Physics: mu = 1, omega = 0,
    theta = 1e-6 deg, pmo = 1, CFL = 0.4, v in [-1, 1]; periodic unless
    the name ends in _open."""
    base = dict(CFL=0.4, flags=VAC | COLL, omega=0.0, theta=1e-6, pmo=1.0, mu=1.0,
                mode=mode, scheme=scheme)
    if name == "c0":   # nz=1024 on [-600,600], nvz=64, 2000 steps
        return Workload(name, RunParams(nz=1024, nvz=64, z0=-600, z1=600, **base), 2000,
                        0.6, 0.5, 0.9)
    if name == "c1":   # nz=65536 on [-1500,1500], nvz=128, 20000 steps, FV and FD
        return Workload(name, RunParams(nz=65536, nvz=128, z0=-1500, z1=1500, **base), 20000,
                        0.6, 0.53, 0.92, gaussian_seed=True)
    if name == "c1_open":  # c1 with open boundaries
        return Workload(name, RunParams(nz=65536, nvz=128, z0=-1500, z1=1500, bc=OPEN, **base),
                        20000, 0.6, 0.53, 0.92, gaussian_seed=True)
    if name == "wide":  # nz=16384 on [-375,375] (c1's dz), nvz=512, c1's physics
        return Workload(name, RunParams(nz=16384, nvz=512, z0=-375, z1=375, **base), 20000,
                        0.6, 0.53, 0.92, gaussian_seed=True)
    if name == "wide_open":  # as wide with nvz=480 (rows padded to 512) and open boundaries
        return Workload(name, RunParams(nz=16384, nvz=480, z0=-375, z1=375, bc=OPEN, **base),
                        20000, 0.6, 0.53, 0.92, gaussian_seed=True)
    if name == "c4":   # nz=2^21 on [-20000,20000], nvz=256, 5000 steps
        return Workload(name, RunParams(nz=2 ** 21, nvz=256, z0=-20000, z1=20000, **base), 5000,
                        0.6, 0.5, 0.9)
    raise KeyError(name)
