"""All four RK4 stages per launch on rows of 512 bins (FAST variant 5 on wide
grids: FD with 3rd-order dissipation).

A team of 2 blocks of 8 wavefronts owns a contiguous z range, each block 256
bins of it; a z row's velocity moments, for each of the four stages, are
summed over the team's two blocks through device memory (tagged words, read
two z steps after they were written).  The per-cell arithmetic is the
single-block four-stage kernel's, so the results agree with the oracle and
with STRICT mode (the reference's arithmetic bit for bit) within the FAST
bounds, with the multi-block two-stage kernel (variant 6) normwise, and bit
for bit with themselves whatever the team count, call pattern or graph
replay.  The reference runs any nvz (rhs_fd.hpp:85-126 loops over all bins).
Every synchronisation also checks that no block gave up waiting for its
partner (the library raises if one did).
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


def _quad(nz, nvz, flags=O.VAC | O.COLL, **kw):
    args = dict(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=O.FD, flags=flags,
                omega=0.3, theta=37.0, pmo=1.0, mu=1.0, **kw)
    return O.RunParams(**args), C.RunParams(mode=C.FAST, **args)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


# (flags, boundary, nvz): uniform rows (512: every group of 32 bins of one
# sign of v) with the benchmark physics, the matter term, no advection and
# no collective term; padded rows with a group across v = 0 (480; 449, the
# narrowest row padded to 512: one live bin in its last live group and a
# group of padding only; 511); the open boundary with and without the
# matter term
CASES = [
    (O.VAC | O.COLL, O.PERIODIC, 449),
    (O.VAC | O.COLL, O.OPEN, 511),
    (O.VAC | O.COLL, O.PERIODIC, 512),
    (O.VAC | O.MAT | O.COLL, O.PERIODIC, 512),
    (O.COLL | O.ADVEC_OFF, O.PERIODIC, 512),
    (O.VAC, O.PERIODIC, 512),
    (O.VAC | O.COLL, O.PERIODIC, 480),
    (O.VAC | O.MAT | O.COLL, O.PERIODIC, 480),
    (O.VAC | O.COLL, O.OPEN, 512),
    (O.VAC | O.COLL, O.OPEN, 480),
    (O.VAC | O.MAT | O.COLL, O.OPEN, 480),
]


@pytest.mark.parametrize("flags,bc,nvz", CASES)
def test_quad_wide_rk4_matches_oracle(gpu, flags, bc, nvz):
    po, pc = _quad(64, nvz, flags, bc=bc)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=71, amp=1e-3)
    print(f"quad wide rk4 nvz={nvz} flags={flags} bc={bc}", flush=True)
    ref = O.step(po, s, 13)  # odd: the state ends in the second buffer
    with C.Simulation(pc) as sim:
        info = sim.info()
        assert info["fast_variant"] == 5 and info["range_teams"] > 1
        assert "quad_wide" in sim.stage_kernel_name(0)
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    print(f"  {info['range_teams']} teams: rel {_rel(out, ref):.3e}", flush=True)
    assert np.isfinite(out).all()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("bc", [O.PERIODIC, O.OPEN])
def test_quad_wide_agrees_with_pair_wide_and_strict(gpu, monkeypatch, bc):
    """60 steps of the flavour-unstable regime on nz = 256 x 512: the
    multi-block four-stage kernel against the multi-block two-stage kernel
    and against STRICT mode's per-cell loop."""
    po, pc = _quad(256, 512, bc=bc)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=72, amp=1e-3)
    outs = {}
    for name, mode, variant in (("strict", C.STRICT, "5"), ("v6", C.FAST, "6"),
                                ("v5", C.FAST, "5")):
        monkeypatch.setenv("COSENU_FAST_VARIANT", variant)
        pc.mode = mode
        print(f"quad wide vs others bc={bc}: {name}", flush=True)
        with C.Simulation(pc) as sim:
            if mode == C.FAST:
                assert sim.info()["fast_variant"] == int(variant)
            sim.upload(s)
            sim.step(60)
            outs[name] = sim.download()
    print(f"  v5-v6 {_rel(outs['v5'], outs['v6']):.3e}, v5-strict "
          f"{_rel(outs['v5'], outs['strict']):.3e}", flush=True)
    assert np.isfinite(outs["v5"]).all()
    assert _rel(outs["v5"], outs["v6"]) <= 1e-12
    assert _rel(outs["v5"], outs["strict"]) <= 1e-10


def test_quad_wide_graph_eager_and_call_pattern_identical(gpu, monkeypatch):
    """16-step graphs, eager steps and odd call lengths agree bit for bit: the
    cross-block sums do not depend on which block finishes first."""
    po, pc = _quad(384, 512)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=73, amp=1e-3)
    outs = []
    for gsteps, calls in (("16", [40]), ("0", [40]), ("16", [3, 20, 17]), ("16", [1, 16, 16, 7])):
        monkeypatch.setenv("COSENU_GRAPH_STEPS", gsteps)
        print(f"quad wide graph steps {gsteps}, calls {calls}", flush=True)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            for n in calls:
                sim.step(n)
            outs.append(sim.download())
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


@pytest.mark.parametrize("nvz", [512, 480])
def test_quad_wide_team_counts_identical(gpu, monkeypatch, nvz):
    """One team per resident slot against twice as many teams (more blocks
    than the device holds at once, so later teams form as earlier ones
    finish): bit for bit."""
    po, pc = _quad(1024, nvz)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=74, amp=1e-3)
    outs, teams = [], []
    for ch in ("-1", "-2"):
        monkeypatch.setenv("COSENU_ZCHUNK", ch)
        print(f"quad wide team count x{ch[1:]} nvz={nvz}", flush=True)
        with C.Simulation(pc) as sim:
            teams.append(sim.info()["range_teams"])
            sim.upload(s)
            sim.step(9)
            outs.append(sim.download())
    print(f"  teams {teams}", flush=True)
    assert teams[1] > teams[0]
    assert np.array_equal(outs[0], outs[1])


def test_quad_wide_state_io_and_diagnostics(gpu):
    """After an odd number of steps: downloads, diagnostics (trace drift at
    the FD scheme's level), the asynchronous P_ee snapshot, a new upload and
    a single RHS address the current state."""
    po, pc = _quad(128, 512)
    G, bG, a = seeded_ic(po, 0.6, 0.5, 0.9, seed=75, amp=1e-3)
    _, _, b = seeded_ic(po, 0.7, 0.4, 1.0, seed=76, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(a, C.STATE0)
        sim.upload(a)
        d0 = sim.diagnostics()
        sim.step(41)
        s41 = sim.download()
        d41 = sim.diagnostics()
        buf = C.PinnedBuffer(pc.nz * pc.nvz)
        sim.snapshot_pee_async(buf)
        sim.snapshot_wait()
        pee = buf.array.reshape(pc.nvz, pc.nz).copy()
        buf.free()
        sim.upload(b)
        r = sim.rhs()
        sim.step(3)
        s3 = sim.download()
    assert d41["nonfinite"] == 0
    assert abs(d41["trace"] - d0["trace"]) <= 1e-12 * abs(d0["trace"])
    assert np.allclose(pee, s41[0] / a[0], rtol=1e-13, atol=0)
    assert _rel(s41, O.step(po, a, 41, moment_form=True)) <= 1e-10
    assert _rel(r, O.rhs(po, b)) <= 1e-12
    assert _rel(s3, O.step(po, b, 3)) <= 1e-10


def test_quad_wide_give_up_is_reported(gpu, monkeypatch):
    """The bounded wait's error path, forced from the host: with
    COSENU_WIDE_GIVE_UP_TEST every wait looks for a tag no block writes and
    gives up after 10 us (every kernel still ends normally); the next
    synchronisation reports the error.  A context made afterwards runs
    normally."""
    po, pc = _quad(64, 512)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=77, amp=1e-3)
    monkeypatch.setenv("COSENU_WIDE_GIVE_UP_TEST", "1")
    print("quad wide give-up test: forced", flush=True)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(3)
        with pytest.raises(C.CosenuError, match="waited more than"):
            sim.sync()
        sim.sync()  # the error is reported once; the word is cleared
    monkeypatch.delenv("COSENU_WIDE_GIVE_UP_TEST")
    print("quad wide give-up test: normal context", flush=True)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(3)
        out = sim.download()
    assert _rel(out, O.step(po, s, 3)) <= 1e-10


@pytest.mark.parametrize("nz,bc", [(16, O.PERIODIC), (100, O.PERIODIC), (100, O.OPEN)])
def test_quad_wide_small_and_ragged_ranges(gpu, nz, bc):
    """The smallest grid the kernel takes (nz = 16: teams of 2 rows, ranges
    shorter than a stage's halo) and a z extent that splits into ranges of
    unequal length: within 1e-10 of the oracle after 13 steps."""
    po, pc = _quad(nz, 512, bc=bc)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=78, amp=1e-3)
    print(f"quad wide small/ragged nz={nz} bc={bc}", flush=True)
    ref = O.step(po, s, 13)
    with C.Simulation(pc) as sim:
        info = sim.info()
        assert info["fast_variant"] == 5 and info["range_teams"] > 1
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    print(f"  {info['range_teams']} teams: rel {_rel(out, ref):.3e}", flush=True)
    assert np.isfinite(out).all()
    assert _rel(out, ref) <= 1e-10


def test_quad_wide_autotune_times_the_wide_kernels(gpu, monkeypatch):
    """The autotuner times the barrier-free kernel and both multi-block
    kernels on rows of 512 bins (FD) and records the fastest."""
    monkeypatch.delenv("COSENU_FAST_VARIANT")
    variant, chunk, ms = C.autotune(C.RunParams(nz=1024, nvz=512, scheme=C.FD, mode=C.FAST))
    print(f"autotune FD nz=1024 nvz=512: variant {variant}, chunk {chunk}, {ms:.3f} ms", flush=True)
    assert variant in (3, 5, 6) and ms > 0
    assert variant == 3 or chunk < 0  # the multi-block kernels run contiguous ranges only


@pytest.mark.parametrize("forced,scheme,flags,nvz,expect", [
    ("5", O.FD, O.VAC | O.COLL, 512, 5),
    ("5", O.FV, O.VAC | O.COLL, 512, 6),            # no four-stage FV: two stages
    ("5", O.FD, O.VAC | O.COLL | O.KO5, 512, 6),    # nor with 5th-order dissipation
    ("5", O.FD, O.VAC | O.COLL, 1024, 6),           # nor on rows of 1024 bins
    ("4", O.FV, O.VAC | O.COLL, 512, 3),            # single-block kernels: barrier-free
    ("6", O.FD, O.VAC | O.COLL, 300, 3),            # rows of 320 bins: barrier-free
])
def test_forced_variant_falls_back_on_wide_rows(gpu, monkeypatch, forced, scheme, flags, nvz,
                                                expect):
    """COSENU_FAST_VARIANT on rows wider than 256 bins: a kernel that does not
    take the grid falls back as on narrower rows (four stages to two, then
    the barrier-free kernel)."""
    monkeypatch.setenv("COSENU_FAST_VARIANT", forced)
    pc = C.RunParams(nz=64, nvz=nvz, scheme=scheme, flags=flags, mode=C.FAST)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == expect


def test_two_contexts_interleaved_match_solo_runs(gpu):
    """Two contexts of different wide grids stepped in turns on one device
    (each with its own exchange rings and ticket counter): every state
    bit-identical to the same run stepped alone."""
    po1, pc1 = _quad(256, 512)
    po2, pc2 = _quad(384, 480, bc=O.OPEN)
    _, _, s1 = seeded_ic(po1, 0.6, 0.5, 0.9, seed=91, amp=1e-3)
    _, _, s2 = seeded_ic(po2, 0.7, 0.4, 1.0, seed=92, amp=1e-3)
    solo = []
    for pc, s in ((pc1, s1), (pc2, s2)):
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(25)
            solo.append(sim.download())
    print("two wide contexts interleaved", flush=True)
    with C.Simulation(pc1) as a, C.Simulation(pc2) as b:
        a.upload(s1)
        b.upload(s2)
        for n in (3, 7, 1, 9, 5):
            a.step(n)
            b.step(n)
        outs = [a.download(), b.download()]
    assert np.array_equal(outs[0], solo[0]) and np.array_equal(outs[1], solo[1])
