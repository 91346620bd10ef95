"""The full nz=1024 x nvz=64 FV workload, 2000 RK4 steps, on the GPU against
the golden vector written by tests/golden/make_c0_fixture.py from the
reference's own build (oracle/_ref).

STRICT: the SHA-256 of the final state equals the reference's (bit-identical
after 2000 nonlinear steps).  FAST: P_ee agrees normwise within 1e-10
(||dP||_inf / ||P||_inf on every 4th z cell, all bins) and the total trace,
lepton number and <P_ee> within 1e-10 relative.
"""
import hashlib
import os

import numpy as np
import pytest

import cosenu_amd as C
from cosenu_amd.initial import workload

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c0_fv_2000.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(FIX, allow_pickle=False)


def _run(mode):
    wl = workload("c0", mode=mode)
    G, bG, s0 = wl.ic()
    with C.Simulation(wl.params) as sim:
        sim.set_profile(G, bG)
        sim.upload(s0, C.STATE0)
        sim.upload(s0)
        sim.step(wl.steps)
        out = sim.download()
        d = sim.diagnostics()
    return s0, out, d


def test_c0_initial_state_matches_fixture(golden):
    wl = workload("c0")
    _, _, s0 = wl.ic()
    assert hashlib.sha256(np.ascontiguousarray(s0).tobytes()).hexdigest() == str(golden["sha256_state0"])


@pytest.mark.gpu
def test_c0_strict_bit_identical_after_2000_steps(gpu, golden):
    s0, out, d = _run(C.STRICT)
    assert hashlib.sha256(np.ascontiguousarray(out).tobytes()).hexdigest() == str(golden["sha256"])


@pytest.mark.gpu
def test_c0_fast_normwise_after_2000_steps(gpu, golden):
    s0, out, d = _run(C.FAST)
    pee = (out[0] / s0[0])[:, ::int(golden["z_stride"])]
    ref = golden["pee_sub"]
    linf = np.max(np.abs(pee - ref)) / np.max(np.abs(ref))
    l2 = np.sqrt(np.mean((pee - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    print(f"c0 FAST vs reference after 2000 steps: P_ee normwise Linf {linf:.3e}, L2 {l2:.3e}")
    assert linf <= 1e-10 and l2 <= 1e-10
    names = [str(n) for n in golden["diag_names"]]
    gd = dict(zip(names, golden["diag"].tolist()))
    for k in ("trace", "btrace", "lepton", "Pee", "Pbee", "M0_norm"):
        assert d[k] == pytest.approx(gd[k], rel=1e-10), k
    assert d["nonfinite"] == 0
