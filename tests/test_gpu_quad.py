"""All four RK4 stages in one launch (FAST variant 5, FD scheme with 3rd-order
dissipation): a team marches along z with the outputs of stages 1-3 in
register windows, each stage two cells behind the previous one, and the
increment sum Q = (Y2 - y) + 2 (Y3 - y) in LDS.  The state buffers alternate
every step.

The stage arithmetic is that of the two-stage kernel (variant 4), so the
results agree with it normwise (only the compiler's contraction of the final
combination may differ), and with the reference arithmetic (the oracle)
within the FAST bounds.  FV and 5th-order dissipation fall back to
variant 4.
"""
import numpy as np
import pytest

import oracle as O
import cosenu_amd as C
from cosenu_amd.initial import seeded_ic

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def variant5(monkeypatch):
    monkeypatch.setenv("COSENU_FAST_VARIANT", "5")


def _quad(nz, nvz, flags=O.VAC | O.COLL, scheme=O.FD, **kw):
    args = dict(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=scheme, flags=flags,
                omega=0.3, theta=37.0, pmo=1.0, mu=1.0, **kw)
    return O.RunParams(**args), C.RunParams(mode=C.FAST, **args)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("nvz", [40, 64, 128, 256])
@pytest.mark.parametrize("flags", [O.VAC | O.COLL, O.VAC | O.MAT | O.COLL, O.COLL,
                                   O.COLL | O.ADVEC_OFF, O.VAC])
def test_quad_rk4_matches_oracle(gpu, flags, nvz):
    po, pc = _quad(96, nvz, flags)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=41, amp=1e-3)
    ref = O.step(po, s, 13)  # odd: the state ends in the second buffer
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 5
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("nz", [16, 17, 23, 250])
def test_quad_small_and_ragged_grids(gpu, nz):
    """The smallest grid the kernel takes (nz = 16: the 6-row halo of stage 1
    wraps once) and ranges of uneven length."""
    po, pc = _quad(nz, 64, O.VAC | O.MAT | O.COLL)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=42, amp=1e-3)
    ref = O.step(po, s, 7)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 5
        sim.upload(s)
        sim.step(7)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


def test_quad_agrees_with_two_stage_kernel(gpu, monkeypatch):
    po, pc = _quad(512, 128, O.VAC | O.MAT | O.COLL)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=43, amp=1e-3)
    outs = {}
    for v in ("4", "5"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", v)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(v)
            sim.upload(s)
            sim.step(41)
            outs[v] = sim.download()
    assert _rel(outs["5"], outs["4"]) <= 1e-12


def test_quad_graph_eager_and_call_pattern_identical(gpu, monkeypatch):
    """16-step graphs, eager steps and odd call lengths (which move the
    current state to the other buffer between graph replays) agree bit for
    bit."""
    po, pc = _quad(384, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=44, amp=1e-3)
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


@pytest.mark.parametrize("nvz,chunk", [(128, "24"), (128, "15"), (128, "2"), (128, "1"),
                                       (128, "-1"), (128, "-2"), (40, "15"), (40, "-1"),
                                       (256, "15"), (256, "-1")])
def test_quad_launch_shapes_identical(gpu, monkeypatch, nvz, chunk):
    """Chunks of 48 rows against chunks of 24, 15, 2 and 1 rows and contiguous
    ranges: bit for bit.  A cell falls in a different one of the z loop's 5
    unrolled steps depending on where its team's range starts; every copy
    rounds it alike (padding bins: nvz = 40; 8 wavefronts per team:
    nvz = 256)."""
    po, pc = _quad(480, nvz)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=45, amp=1e-3)
    outs = []
    for ch in ("48", chunk):
        monkeypatch.setenv("COSENU_ZCHUNK", ch)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(9)
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1])


def test_quad_state_io_between_odd_calls(gpu):
    """Uploads, downloads, diagnostics and the RHS evaluation address the
    current state buffer after an odd number of steps."""
    po, pc = _quad(256, 64)
    G, bG, a = seeded_ic(po, 0.6, 0.5, 0.9, seed=46, amp=1e-3)
    _, _, b = seeded_ic(po, 0.7, 0.4, 1.0, seed=47, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(a, C.STATE0)
        sim.upload(a)
        sim.step(5)
        s5 = sim.download()
        d5 = sim.diagnostics()
        k5 = sim.rhs()
        sim.upload(b)
        sim.step(7)
        got = sim.download()
    ref5 = O.step(po, a, 5)
    assert _rel(s5, ref5) <= 1e-10
    dref = O.diagnostics(po, G, bG, a, ref5)
    assert d5["trace"] == pytest.approx(dref["trace"], rel=1e-12)
    kref = O.rhs(po, ref5)
    assert _rel(k5, kref) <= 1e-10
    with C.Simulation(pc) as sim:
        sim.upload(b)
        sim.step(7)
        want = sim.download()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("scheme,flags,want", [(O.FV, O.VAC | O.COLL, 4),
                                               (O.FD, O.COLL | O.KO5, 4)])
def test_quad_unsupported_shapes_fall_back(gpu, scheme, flags, want):
    """FV and 5th-order dissipation read z +- 3 per stage: the forced variant
    falls back to the two-stage kernel, whose results match the oracle."""
    po, pc = _quad(128, 64, flags, scheme=scheme)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=48, amp=1e-3)
    ref = O.step(po, s, 5)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == want
        sim.upload(s)
        sim.step(5)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


def test_quad_rk4_combination_has_no_drift(gpu):
    """With a zero right-hand side (no advection, vacuum or collective term,
    no dissipation) RK4 leaves the state unchanged: the increment form of
    the combination keeps it exactly."""
    po, pc = _quad(256, 64, O.ADVEC_OFF, ko=0.0)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=49, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(2000)
        out = sim.download()
    drift = abs(out.sum() - s.sum()) / abs(s.sum())
    assert drift < 1e-14
    assert _rel(out, s) < 1e-13


def test_quad_autotune_considers_four_stage_kernel(gpu, tmp_path, monkeypatch):
    table = tmp_path / "tune.txt"
    monkeypatch.setenv("COSENU_TUNE_FILE", str(table))
    monkeypatch.delenv("COSENU_FAST_VARIANT", raising=False)
    # (a shape no other test uses: the tuned entry stays in the process's table)
    p = C.RunParams(nz=3072, nvz=128, scheme=C.FD, mode=C.FAST)
    v, L, ms = C.autotune(p)
    assert v in (1, 2, 3, 4, 5) and ms > 0
    with C.Simulation(p) as sim:
        assert sim.info()["fast_variant"] == v
    po = O.RunParams(nz=3072, nvz=128, scheme=O.FD)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=50, amp=1e-3)
    outs = {}
    for var in ("4", "5"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", var)
        monkeypatch.setenv("COSENU_ZCHUNK", "-1")
        with C.Simulation(p) as sim:
            sim.upload(s)
            sim.step(4)
            outs[var] = sim.download()
    assert _rel(outs["5"], outs["4"]) <= 1e-12


def test_quad_is_the_fd_default_and_times_one_launch(gpu, monkeypatch):
    """Without a forced variant FD runs of a size that fills the device take
    the four-stage kernel; timing reports one launch per step."""
    monkeypatch.delenv("COSENU_FAST_VARIANT", raising=False)
    monkeypatch.delenv("COSENU_ZCHUNK", raising=False)
    monkeypatch.delenv("COSENU_TUNE_FILE", raising=False)
    po, pc = _quad(8192, 128)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=51, amp=1e-3)
    with C.Simulation(pc) as sim:
        info = sim.info()
        assert info["fast_variant"] == 5 and info["range_teams"] > 0
        sim.upload(s)
        sim.set_timing(True)
        sim.step(5)
        ms, cnt = sim.timing()
        names = [sim.stage_kernel_name(k) for k in range(4)]
        out = sim.download()
    assert cnt == [5, 0, 0, 0] and ms[0] > 0
    assert "quad" in names[0] and not any(names[1:])
    assert _rel(out, O.step(po, s, 5, moment_form=True, threads=8)) <= 1e-10


@pytest.mark.parametrize("chunk", ["", "48"])
def test_quad_batch_bit_identical_to_solo_runs(gpu, monkeypatch, chunk):
    """A batch of FD runs of uneven nz and nvz (groups of nvz 64, 128, 256,
    one with padding bins) takes the four-stage kernel for every group and
    leaves each run bit-identical to the run advanced alone; chunks of 48
    rows leave ragged last items."""
    monkeypatch.delenv("COSENU_BATCH_VARIANT", raising=False)
    if chunk:
        monkeypatch.setenv("COSENU_BATCH_ZCHUNK", chunk)
    specs = [(256, 64), (100, 40), (320, 128), (48, 256), (17, 64), (200, 128)]
    rng = np.random.default_rng(7)
    runs, ics = [], []
    for k, (nz, nvz) in enumerate(specs):
        p = C.RunParams(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=C.FD,
                        flags=C.VAC | C.COLL | (C.MAT if k % 2 else 0),
                        omega=float(rng.uniform(0, 0.5)), theta=float(rng.uniform(1, 40)),
                        mu=float(rng.uniform(0.5, 1.5)), mode=C.FAST)
        runs.append(p)
        ics.append(seeded_ic(p, rng.uniform(0.4, 0.8), rng.uniform(0.4, 0.8),
                             rng.uniform(0.7, 1.1), seed=200 + k, amp=1e-3))
    with C.Batch(runs) as b:
        assert b.info()["fast_variant"] == 5
        for r, (G, bG, s0) in enumerate(ics):
            b.set_profile(r, G, bG)
            b.upload(r, s0)
        b.step(21)
        outs = [b.download(r) for r in range(len(runs))]
    for r, (p, (G, bG, s0)) in enumerate(zip(runs, ics)):
        with C.Simulation(p) as sim:
            assert sim.info()["fast_variant"] == 5
            sim.upload(s0)
            sim.step(21)
            solo = sim.download()
        assert np.array_equal(outs[r], solo), r
    p, (_, _, s0) = runs[1], ics[1]
    po = O.RunParams(nz=p.nz, nvz=p.nvz, z0=p.z0, z1=p.z1, scheme=O.FD, flags=p.flags,
                     omega=p.omega, theta=p.theta, mu=p.mu)
    assert _rel(outs[1], O.step(po, s0, 21)) <= 1e-10


# The open boundary (nuosc.hpp:247-277): y rows below the grid read row 0,
# rows above it extrapolate linearly; each stage's output follows the same
# rule at its ghost rows.
def _open_state(po, seed):
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=seed, amp=1e-3)
    # a z-dependent background so that the boundary rules matter
    return s * (1.0 + 0.05 * np.sin(np.linspace(0.0, 3.0, po.nz)))[None, None, :]


@pytest.mark.parametrize("nvz", [40, 64, 128, 256])
@pytest.mark.parametrize("flags", [O.VAC | O.COLL, O.VAC | O.MAT | O.COLL])
def test_quad_open_boundary_matches_oracle(gpu, flags, nvz):
    po, pc = _quad(96, nvz, flags, bc=O.OPEN)
    s = _open_state(po, 51)
    ref = O.step(po, s, 13)
    print(f"quad open flags={flags} nvz={nvz}", flush=True)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 5
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("nz", [16, 17, 23])
def test_quad_open_boundary_small_grids(gpu, nz):
    """Grids where every team's halo reaches both boundaries."""
    po, pc = _quad(nz, 64, O.VAC | O.MAT | O.COLL, bc=O.OPEN)
    s = _open_state(po, 52)
    ref = O.step(po, s, 7)
    print(f"quad open nz={nz}", flush=True)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 5
        sim.upload(s)
        sim.step(7)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("chunk", ["2", "1", "-1", "-2"])
def test_quad_open_boundary_launch_shapes_identical(gpu, monkeypatch, chunk):
    po, pc = _quad(96, 64, bc=O.OPEN)
    s = _open_state(po, 53)
    outs = []
    for ch in ("48", chunk):
        monkeypatch.setenv("COSENU_ZCHUNK", ch)
        print(f"quad open chunk={ch}", flush=True)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(5)
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1])
    assert _rel(outs[0], O.step(po, s, 5)) <= 1e-10


def test_quad_open_boundary_agrees_with_two_stage_kernel(gpu, monkeypatch):
    po, pc = _quad(512, 128, O.VAC | O.MAT | O.COLL, bc=O.OPEN)
    s = _open_state(po, 54)
    outs = {}
    for v in ("4", "5"):
        monkeypatch.setenv("COSENU_FAST_VARIANT", v)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(v)
            sim.upload(s)
            sim.step(41)
            outs[v] = sim.download()
    assert _rel(outs["5"], outs["4"]) <= 1e-12
