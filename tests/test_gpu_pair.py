"""Two RK4 stages per launch (FAST variant 4): stages 1+2 and 3+4 each run as
one sweep with the intermediate stage state kept in registers, and the state
buffers alternate every step.

The arithmetic per cell is the split-lane kernel's, so the results agree with
variant 2 normwise (the operation order of the second stage's moments is the
same; only which lanes compute which halo cells differs) and with the
reference arithmetic (the oracle) within the FAST bounds.
"""
import numpy as np
import pytest

import oracle as O
import cosenu_amd as C
from cosenu_amd.initial import seeded_ic

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def variant4(monkeypatch):
    monkeypatch.setenv("COSENU_FAST_VARIANT", "4")


def _pair(scheme, nz, nvz, flags=O.VAC | O.COLL, **kw):
    args = dict(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=scheme, flags=flags,
                omega=0.3, theta=37.0, pmo=1.0, mu=1.0, **kw)
    return O.RunParams(**args), C.RunParams(mode=C.FAST, **args)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("nvz", [40, 64, 128, 256])
@pytest.mark.parametrize("scheme,flags", [(O.FV, O.VAC | O.COLL), (O.FV, O.VAC | O.MAT | O.COLL),
                                          (O.FD, O.VAC | O.COLL), (O.FD, O.COLL | O.KO5),
                                          (O.FV, O.COLL | O.ADVEC_OFF)])
def test_pair_rk4_matches_oracle(gpu, scheme, flags, nvz):
    po, pc = _pair(scheme, 96, nvz, flags)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=31, amp=1e-3)
    ref = O.step(po, s, 13)  # odd: the state ends in the second buffer
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 4
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


def test_pair_agrees_with_split_lane_kernel(gpu, monkeypatch):
    po, pc = _pair(O.FV, 512, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=32, amp=1e-3)
    outs = {}
    for v in ("2", "4"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", v)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(v)
            sim.upload(s)
            sim.step(41)
            outs[v] = sim.download()
    assert _rel(outs["4"], outs["2"]) <= 1e-12


def test_pair_graph_eager_and_call_pattern_identical(gpu, monkeypatch):
    """16-step graphs, eager steps and odd call lengths (which move the
    current state to the other buffer between graph replays) agree bit for
    bit."""
    po, pc = _pair(O.FV, 384, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=33, amp=1e-3)
    outs = []
    for gsteps, calls in (("16", [40]), ("0", [40]), ("16", [3, 20, 17]), ("16", [1, 16, 16, 7])):
        monkeypatch.setenv("COSENU_GRAPH_STEPS", gsteps)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            for n in calls:
                sim.step(n)
            outs.append(sim.download())
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


@pytest.mark.parametrize("chunk", ["24", "15", "2", "1", "-1", "-2"])
@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_pair_launch_shapes_identical(gpu, monkeypatch, scheme, chunk):
    """Chunks of 48 rows against chunks of 24, 15, 2 and 1 rows and contiguous
    ranges: bit for bit.  (FD runs two z steps per loop trip; chunks of 48
    rows give an odd step count, which ends with an idle step, 15 rows an
    even one.)"""
    po, pc = _pair(scheme, 480, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=34, amp=1e-3)
    outs = []
    for ch in ("48", chunk):
        monkeypatch.setenv("COSENU_ZCHUNK", ch)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(9)
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1])


def test_pair_state_io_between_odd_calls(gpu):
    """Uploads, downloads, diagnostics and snapshots address the current
    state buffer after an odd number of steps."""
    po, pc = _pair(O.FV, 256, 64)
    G, bG, a = seeded_ic(po, 0.6, 0.5, 0.9, seed=35, amp=1e-3)
    _, _, b = seeded_ic(po, 0.7, 0.4, 1.0, seed=36, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(a, C.STATE0)
        sim.upload(a)
        sim.step(5)
        s5 = sim.download()
        d5 = sim.diagnostics()
        buf = C.PinnedBuffer(pc.nz * pc.nvz)
        sim.snapshot_pee_async(buf)
        sim.snapshot_wait()
        pee = buf.array.reshape(pc.nvz, pc.nz).copy()
        buf.free()
        sim.upload(b)
        sim.step(7)
        got = sim.download()
    ref5 = O.step(po, a, 5)
    assert _rel(s5, ref5) <= 1e-10
    dref = O.diagnostics(po, G, bG, a, ref5)
    assert d5["trace"] == pytest.approx(dref["trace"], rel=1e-12)
    assert np.allclose(pee, s5[0] / a[0], rtol=1e-13, atol=0)
    with C.Simulation(pc) as sim:
        sim.upload(b)
        sim.step(7)
        want = sim.download()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_pair_small_problem_default_launch(gpu, scheme):
    """A problem too small to fill the device (nz = 1024, nvz = 64, the
    CPU-sized case's grid) gets ranges of 2 rows, one wavefront per SIMD;
    its steps match the oracle."""
    po, pc = _pair(scheme, 1024, 64)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=37, amp=1e-3)
    ref = O.step(po, s, 6)
    with C.Simulation(pc) as sim:
        info = sim.info()
        assert info["fast_variant"] == 4
        assert info["range_teams"] == 512 and info["z_chunk"] == 2
        sim.upload(s)
        sim.step(6)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


def test_pair_rk4_combination_has_no_drift(gpu):
    """With a zero right-hand side (no advection, vacuum or collective term)
    RK4 leaves the state unchanged.  The two-stage kernel's combination must
    not round state-sized products: y' = (Y2 + 2 Y3 - y)/3 + Y4/3 with the
    rounded 1/3 shrinks every value by ~5.6e-17 per step (1.1e-13 after
    2000 steps); the increment form y + ((Q + (Y4 - y))/3 + (dt/6) K4) keeps
    it exactly."""
    po, pc = _pair(O.FV, 256, 64, O.ADVEC_OFF)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=38, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(2000)
        out = sim.download()
    drift = abs(out.sum() - s.sum()) / abs(s.sum())
    assert drift < 1e-14
    assert _rel(out, s) < 1e-13


# The open boundary (nuosc.hpp:247-277): stage-input rows below the grid read
# row 0, rows above it extrapolate linearly; the first stage's output follows
# the same rule at its ghost rows, and the flux ghosts of the two boundary
# cells follow the reference's (rflux[-1] = rflux[0], lflux[nz] extrapolated).
def _open_state(po, seed):
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=seed, amp=1e-3)
    # a z-dependent background so that the boundary rules matter
    return s * (1.0 + 0.05 * np.sin(np.linspace(0.0, 3.0, po.nz)))[None, None, :]


@pytest.mark.parametrize("scheme,flags,nvz", [
    (O.FV, O.VAC | O.COLL, 64), (O.FV, O.VAC | O.MAT | O.COLL, 128), (O.FV, O.VAC | O.COLL, 100),
    (O.FD, O.VAC | O.COLL, 128), (O.FD, O.COLL | O.KO5, 256), (O.FD, O.VAC | O.COLL, 40)])
def test_pair_open_boundary_matches_oracle(gpu, scheme, flags, nvz):
    po, pc = _pair(scheme, 96, nvz, flags, bc=O.OPEN)
    s = _open_state(po, 39)
    ref = O.step(po, s, 13)
    print(f"pair open scheme={scheme} flags={flags} nvz={nvz}", flush=True)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 4
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("chunk", ["2", "1", "-1", "-2"])
@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_pair_open_boundary_launch_shapes_identical(gpu, monkeypatch, scheme, chunk):
    """Ranges of 1 and 2 rows put the boundary cells in teams whose first
    stage covers only ghost rows on one side: bit for bit the results of
    chunks of 48 rows."""
    po, pc = _pair(scheme, 96, 64, bc=O.OPEN)
    s = _open_state(po, 40)
    outs = []
    for ch in ("48", chunk):
        monkeypatch.setenv("COSENU_ZCHUNK", ch)
        print(f"pair open scheme={scheme} chunk={ch}", flush=True)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(5)
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1])
    assert _rel(outs[0], O.step(po, s, 5)) <= 1e-10


def test_pair_open_boundary_agrees_with_split_lane_kernel(gpu, monkeypatch):
    po, pc = _pair(O.FV, 512, 128, bc=O.OPEN)
    s = _open_state(po, 41)
    outs = {}
    for v in ("2", "4"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", v)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(v)
            sim.upload(s)
            sim.step(41)
            outs[v] = sim.download()
    assert _rel(outs["4"], outs["2"]) <= 1e-12
