"""The barrier-free FAST kernel (variant 3): every wavefront marches its own
32 velocity bins in the upwind direction and reads the per-z moments of its
input from the launch that wrote it.

It is the default for small problems (at most 2 wavefronts per SIMD in all)
and is selected with COSENU_FAST_VARIANT=3 (or by the autotuner) where it
applies (nvz a multiple of 64, periodic boundaries, every group of 32 bins of
one sign of v); elsewhere the split-lane kernel (variant 2) runs.  Both
evaluate the same mathematics with different summation orders, so they agree
normwise, and both agree with the reference arithmetic (the oracle) within
the FAST bounds.
"""
import numpy as np
import pytest

import oracle as O
import cosenu_amd as C
from cosenu_amd.initial import seeded_ic

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def variant3(monkeypatch):
    monkeypatch.setenv("COSENU_FAST_VARIANT", "3")


def _pair(scheme, nz, nvz, flags=O.VAC | O.COLL, v0=-1.0, v1=1.0, **kw):
    args = dict(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=scheme, flags=flags,
                omega=0.3, theta=37.0, pmo=1.0, mu=1.0, v0=v0, v1=v1, **kw)
    return O.RunParams(**args), C.RunParams(mode=C.FAST, **args)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("nvz", [64, 128, 256])
@pytest.mark.parametrize("scheme,flags", [(O.FV, O.VAC | O.COLL), (O.FV, O.VAC | O.MAT | O.COLL),
                                          (O.FD, O.VAC | O.COLL), (O.FD, O.COLL | O.KO5),
                                          (O.FV, O.COLL | O.ADVEC_OFF)])
def test_stream_rk4_matches_oracle(gpu, scheme, flags, nvz):
    po, pc = _pair(scheme, 96, nvz, flags)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=21, amp=1e-3)
    ref = O.step(po, s, 12)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 3
        sim.upload(s)
        sim.step(12)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_stream_rhs_matches_oracle(gpu, scheme):
    po, pc = _pair(scheme, 64, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=22, amp=1e-3)
    ref = O.rhs(po, s)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 3
        sim.upload(s)
        out = sim.rhs()
    assert _rel(out, ref) <= 1e-12


def test_stream_agrees_with_split_lane_kernel(gpu, monkeypatch):
    po, pc = _pair(O.FV, 512, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=23, amp=1e-3)
    outs = {}
    for v in ("2", "3"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", v)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(v)
            sim.upload(s)
            sim.step(40)
            outs[v] = sim.download()
    assert _rel(outs["3"], outs["2"]) <= 1e-12


def test_stream_state_replaced_between_calls(gpu):
    """The moments of the state are rebuilt at every step() call, so a state
    uploaded after stepping advances exactly like a fresh one."""
    po, pc = _pair(O.FV, 256, 128)
    _, _, a = seeded_ic(po, 0.6, 0.5, 0.9, seed=24, amp=1e-3)
    _, _, b = seeded_ic(po, 0.7, 0.4, 1.0, seed=25, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.upload(a)
        sim.step(7)
        sim.upload(b)
        sim.step(9)
        got = sim.download()
    with C.Simulation(pc) as sim:
        sim.upload(b)
        sim.step(9)
        want = sim.download()
    assert np.array_equal(got, want)


def test_stream_graph_and_eager_and_split_calls_identical(gpu, monkeypatch):
    po, pc = _pair(O.FV, 384, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=26, amp=1e-3)
    outs = []
    for gsteps, calls in (("16", [40]), ("0", [40]), ("16", [3, 20, 17])):
        monkeypatch.setenv("COSENU_GRAPH_STEPS", gsteps)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            for n in calls:
                sim.step(n)
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_mixed_sign_groups_fall_back_to_split_lane(gpu):
    """v in [-0.5, 1] with 64 bins: the first group of 32 holds both signs, so
    the split-lane kernel runs; the results still match the oracle."""
    po, pc = _pair(O.FV, 64, 64, v0=-0.5, v1=1.0)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=27, amp=1e-3)
    ref = O.step(po, s, 6)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 2
        sim.upload(s)
        sim.step(6)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_small_problem_default_is_barrier_free(gpu, scheme, monkeypatch):
    """Without a tuned entry or override, a problem of at most 2 wavefronts
    per SIMD in all (nz = 1024, nvz = 64: the CPU-sized case's grid) runs the
    barrier-free kernel in chunks of 2 rows; its steps match the oracle."""
    monkeypatch.delenv("COSENU_FAST_VARIANT", raising=False)
    monkeypatch.delenv("COSENU_ZCHUNK", raising=False)
    po, pc = _pair(scheme, 1024, 64)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=23, amp=1e-3)
    ref = O.step(po, s, 6)
    with C.Simulation(pc) as sim:
        info = sim.info()
        assert info["fast_variant"] == 3 and info["z_chunk"] == 2 and info["range_teams"] == 0
        sim.upload(s)
        sim.step(6)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10
