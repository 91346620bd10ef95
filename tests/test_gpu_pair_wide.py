"""Two RK4 stages per launch on rows of 512 and 1024 bins (FAST variant 6).

A team of 2 or 4 blocks of 8 wavefronts owns a contiguous z range, each block
256 bins of it; a z row's velocity moments are summed over the team's blocks
through device memory (tagged words, read four z steps after they were
written), in a tree that every block rounds alike.  The per-cell arithmetic
is the single-block two-stage kernel's, so the results agree with the oracle
and with STRICT mode (the reference's arithmetic bit for bit) within the FAST
bounds, with the barrier-free kernel (variant 3) normwise, and bit for bit
with themselves whatever the team count, call pattern or graph replay.  The
reference runs any nvz (rhs_fv.hpp:188-233, rhs_fd.hpp:85-126 loop over all
bins).  Every synchronisation also checks that no block gave up waiting for
its partners (the library raises if one did).
"""
import numpy as np
import pytest

import oracle as O
import cosenu_amd as C
from cosenu_amd.initial import seeded_ic

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def variant6(monkeypatch):
    monkeypatch.setenv("COSENU_FAST_VARIANT", "6")


def _pair(scheme, nz, nvz, flags=O.VAC | O.COLL, **kw):
    args = dict(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=scheme, flags=flags,
                omega=0.3, theta=37.0, pmo=1.0, mu=1.0, **kw)
    return O.RunParams(**args), C.RunParams(mode=C.FAST, **args)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


# (scheme, flags, boundary, nvz): uniform rows (512, 1024: every group of 32
# bins of one sign of v), padded rows with a group across v = 0 (480, 1000),
# the open boundary, the matter term, 5th-order dissipation, no advection
CASES = [
    (O.FV, O.VAC | O.COLL, O.PERIODIC, 512),
    (O.FD, O.VAC | O.COLL, O.PERIODIC, 512),
    (O.FV, O.VAC | O.MAT | O.COLL, O.PERIODIC, 512),
    (O.FD, O.COLL | O.KO5, O.PERIODIC, 512),
    (O.FV, O.COLL | O.ADVEC_OFF, O.PERIODIC, 512),
    (O.FV, O.VAC | O.COLL, O.PERIODIC, 480),
    (O.FV, O.VAC | O.COLL, O.OPEN, 512),
    (O.FD, O.VAC | O.COLL, O.OPEN, 480),
    (O.FV, O.VAC | O.COLL, O.PERIODIC, 1024),
    (O.FD, O.VAC | O.COLL, O.PERIODIC, 1024),
    (O.FV, O.VAC | O.MAT | O.COLL, O.OPEN, 1000),
    # the narrowest rows padded to 512 and 1024 bins: one live bin in the
    # last live group, whole groups of padding
    (O.FV, O.VAC | O.COLL, O.PERIODIC, 449),
    (O.FD, O.VAC | O.COLL, O.OPEN, 961),
]


@pytest.mark.parametrize("scheme,flags,bc,nvz", CASES)
def test_pair_wide_rk4_matches_oracle(gpu, scheme, flags, bc, nvz):
    nz = 64 if nvz <= 512 else 32
    po, pc = _pair(scheme, nz, nvz, flags, bc=bc)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=61, amp=1e-3)
    print(f"pair wide rk4 nvz={nvz} scheme={scheme} flags={flags} bc={bc}", flush=True)
    ref = O.step(po, s, 13)  # odd: the state ends in the second buffer
    with C.Simulation(pc) as sim:
        info = sim.info()
        assert info["fast_variant"] == 6 and info["range_teams"] > 1
        assert "pair_wide" in sim.stage_kernel_name(0)
        sim.upload(s)
        sim.step(13)
        out = sim.download()
    print(f"  {info['range_teams']} teams: rel {_rel(out, ref):.3e}", flush=True)
    assert np.isfinite(out).all()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_pair_wide_agrees_with_barrier_free_and_strict(gpu, monkeypatch, scheme):
    """60 steps of the flavour-unstable regime on nz = 256 x 512: the
    multi-block two-stage kernel against the barrier-free kernel and against
    STRICT mode's per-cell loop."""
    po, pc = _pair(scheme, 256, 512)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=62, amp=1e-3)
    outs = {}
    for name, mode, variant in (("strict", C.STRICT, "6"), ("v3", C.FAST, "3"),
                                ("v6", C.FAST, "6")):
        monkeypatch.setenv("COSENU_FAST_VARIANT", variant)
        pc.mode = mode
        print(f"pair wide vs others scheme={scheme}: {name}", flush=True)
        with C.Simulation(pc) as sim:
            if mode == C.FAST:
                assert sim.info()["fast_variant"] == int(variant)
            sim.upload(s)
            sim.step(60)
            outs[name] = sim.download()
    print(f"  v6-v3 {_rel(outs['v6'], outs['v3']):.3e}, v6-strict "
          f"{_rel(outs['v6'], outs['strict']):.3e}", flush=True)
    assert np.isfinite(outs["v6"]).all()
    assert _rel(outs["v6"], outs["v3"]) <= 1e-12
    assert _rel(outs["v6"], outs["strict"]) <= 1e-10


def test_pair_wide_graph_eager_and_call_pattern_identical(gpu, monkeypatch):
    """16-step graphs, eager steps and odd call lengths agree bit for bit: the
    cross-block sums do not depend on which block finishes first."""
    po, pc = _pair(O.FV, 384, 512)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=63, amp=1e-3)
    outs = []
    for gsteps, calls in (("16", [40]), ("0", [40]), ("16", [3, 20, 17]), ("16", [1, 16, 16, 7])):
        monkeypatch.setenv("COSENU_GRAPH_STEPS", gsteps)
        print(f"pair wide graph steps {gsteps}, calls {calls}", flush=True)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            for n in calls:
                sim.step(n)
            outs.append(sim.download())
    for o in outs[1:]:
        assert np.array_equal(outs[0], o)


@pytest.mark.parametrize("scheme,nvz", [(O.FV, 512), (O.FD, 512), (O.FV, 1024)])
def test_pair_wide_team_counts_identical(gpu, monkeypatch, scheme, nvz):
    """One team per resident slot against twice as many teams (more blocks
    than the device holds at once, so later teams form as earlier ones
    finish): bit for bit."""
    po, pc = _pair(scheme, 1024, nvz)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=64, amp=1e-3)
    outs, teams = [], []
    for ch in ("-1", "-2"):
        monkeypatch.setenv("COSENU_ZCHUNK", ch)
        print(f"pair wide team count x{ch[1:]} scheme={scheme} nvz={nvz}", flush=True)
        with C.Simulation(pc) as sim:
            teams.append(sim.info()["range_teams"])
            sim.upload(s)
            sim.step(9)
            outs.append(sim.download())
    print(f"  teams {teams}", flush=True)
    assert teams[1] > teams[0]
    assert np.array_equal(outs[0], outs[1])


def test_pair_wide_state_io_and_diagnostics(gpu):
    """After an odd number of steps: downloads, diagnostics (trace conserved
    on the periodic FV grid), the asynchronous P_ee snapshot, a new upload
    and a single RHS address the current state."""
    po, pc = _pair(O.FV, 128, 512)
    G, bG, a = seeded_ic(po, 0.6, 0.5, 0.9, seed=65, amp=1e-3)
    _, _, b = seeded_ic(po, 0.7, 0.4, 1.0, seed=66, amp=1e-3)
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
    assert abs(d41["btrace"] - d0["btrace"]) <= 1e-12 * abs(d0["btrace"])
    assert np.allclose(pee, s41[0] / a[0], rtol=1e-13, atol=0)
    assert _rel(s41, O.step(po, a, 41, moment_form=True)) <= 1e-10
    assert _rel(r, O.rhs(po, b)) <= 1e-12
    assert _rel(s3, O.step(po, b, 3)) <= 1e-10


def test_pair_wide_give_up_is_reported(gpu, monkeypatch):
    """The bounded wait's error path, forced from the host: with
    COSENU_WIDE_GIVE_UP_TEST every wait looks for a tag no block writes and
    gives up after 10 us (every kernel still ends normally); the next
    synchronisation reports the error.  A context made afterwards runs
    normally."""
    po, pc = _pair(O.FV, 64, 512)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=67, amp=1e-3)
    monkeypatch.setenv("COSENU_WIDE_GIVE_UP_TEST", "1")
    print("pair wide give-up test: forced", flush=True)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(3)
        with pytest.raises(C.CosenuError, match="waited more than"):
            sim.sync()
        sim.sync()  # the error is reported once; the word is cleared
    monkeypatch.delenv("COSENU_WIDE_GIVE_UP_TEST")
    print("pair wide give-up test: normal context", flush=True)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(3)
        out = sim.download()
    assert _rel(out, O.step(po, s, 3)) <= 1e-10
