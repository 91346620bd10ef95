"""FAST mode for velocity grids wider than 256 bins (nvz = 300 .. 1024): the
barrier-free kernel (variant 3).

The single-block team kernels keep the VP/32 wavefronts of a z range in one
block of at most 8; the barrier-free kernel takes every wider grid (and is
the default for rows other than 512 or 1024 bins, which take the multi-block
two-stage kernel, variant 6, tested in test_gpu_pair_wide.py).  Each of its
wavefronts owns 32 bins and the per-z moments pass through memory as one
partial per (z row, group of 32 bins).  Beyond 8 groups a lane first adds
the partials of groups g, g + 8, g + 16, ... and the fixed lane-move tree
then adds the 8 results.  The reference runs any nvz (rhs_fv.hpp:188-233
loops over all bins); these tests pin the wide path against the oracle and
against STRICT mode, which is bit-identical to the reference's arithmetic.
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


def _pair(scheme, nz, nvz, flags=O.VAC | O.COLL, **kw):
    args = dict(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=scheme, flags=flags,
                omega=0.3, theta=37.0, pmo=1.0, mu=1.0, **kw)
    return O.RunParams(**args), C.RunParams(mode=C.FAST, **args)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("nvz", [384, 512, 1024])
@pytest.mark.parametrize("scheme,flags", [(O.FV, O.VAC | O.COLL), (O.FD, O.VAC | O.COLL),
                                          (O.FV, O.VAC | O.MAT | O.COLL), (O.FD, O.COLL | O.KO5)])
def test_wide_rk4_matches_oracle(gpu, scheme, flags, nvz):
    nz = 64 if nvz < 1024 else 32
    po, pc = _pair(scheme, nz, nvz, flags)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=41, amp=1e-3)
    ref = O.step(po, s, 10)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 3
        sim.upload(s)
        sim.step(10)
        out = sim.download()
    print(f"nvz={nvz} scheme={scheme} flags={flags}: rel {_rel(out, ref):.3e}", flush=True)
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("nvz", [512, 1024])
def test_wide_rhs_matches_oracle(gpu, nvz):
    po, pc = _pair(O.FV, 32, nvz)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=42, amp=1e-3)
    ref = O.rhs(po, s)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        out = sim.rhs()
    assert _rel(out, ref) <= 1e-12


@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_wide_fast_agrees_with_strict(gpu, scheme):
    """Longer run, flavour-unstable regime: FAST (moment form, lanes reduced
    across 16 groups) against STRICT (the reference's per-cell loop)."""
    po, pc = _pair(scheme, 256, 512)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=43, amp=1e-3)
    outs = {}
    for mode in (C.STRICT, C.FAST):
        pc.mode = mode
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(60)
            outs[mode] = sim.download()
    assert np.isfinite(outs[C.FAST]).all()
    assert _rel(outs[C.FAST], outs[C.STRICT]) <= 1e-10


def test_wide_diagnostics_conserve_trace(gpu):
    po, pc = _pair(O.FV, 128, 512, flags=O.VAC | O.COLL)
    G, bG, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=44, amp=1e-3)
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(s, C.STATE0)
        sim.upload(s)
        d0 = sim.diagnostics()
        sim.step(40)
        d1 = sim.diagnostics()
    assert d1["nonfinite"] == 0
    assert abs(d1["trace"] - d0["trace"]) <= 1e-12 * abs(d0["trace"])
    assert abs(d1["btrace"] - d0["btrace"]) <= 1e-12 * abs(d0["btrace"])


def test_wide_autotune_times_both_wide_kernels(gpu, monkeypatch):
    """The autotuner times the barrier-free kernel and the multi-block
    two-stage kernel on rows of 512 bins and records the faster."""
    monkeypatch.delenv("COSENU_FAST_VARIANT")
    variant, chunk, ms = C.autotune(C.RunParams(nz=1024, nvz=512, mode=C.FAST))
    print(f"autotune nz=1024 nvz=512: variant {variant}, chunk {chunk}, {ms:.3f} ms", flush=True)
    assert variant in (3, 6) and ms > 0
    assert variant == 3 or chunk < 0  # the two-stage kernel runs contiguous ranges only


# The general form of the barrier-free kernel: bin groups of 32 that straddle
# v = 0 (nvz = 300, 480: the group holding bin nvz/2), padding bins (300 in
# rows of 320, 480 in 512, 1000 in 1024) and the open boundary (nuosc.hpp:247-277:
# ghost rows copy row 0 below the grid and extrapolate linearly above it).
GENERAL = [
    (O.FV, O.VAC | O.COLL, O.PERIODIC, 300),
    (O.FD, O.VAC | O.COLL, O.PERIODIC, 300),
    (O.FV, O.VAC | O.COLL, O.OPEN, 512),
    (O.FD, O.VAC | O.COLL, O.OPEN, 512),
    (O.FV, O.VAC | O.MAT | O.COLL, O.OPEN, 448),
    (O.FD, O.COLL | O.KO5, O.PERIODIC, 480),
    (O.FD, O.VAC | O.MAT | O.COLL | O.KO5, O.OPEN, 480),
    (O.FV, O.VAC | O.COLL, O.PERIODIC, 1000),
]


@pytest.mark.parametrize("scheme,flags,bc,nvz", GENERAL)
def test_general_rk4_matches_oracle(gpu, scheme, flags, bc, nvz):
    nz = 64 if nvz < 1000 else 32
    po, pc = _pair(scheme, nz, nvz, flags, bc=bc)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=45, amp=1e-3)
    ref = O.step(po, s, 10)
    print(f"general rk4 nvz={nvz} scheme={scheme} flags={flags} bc={bc}", flush=True)
    with C.Simulation(pc) as sim:
        assert sim.info()["fast_variant"] == 3
        assert "general" in sim.stage_kernel_name(0)
        sim.upload(s)
        sim.step(10)
        out = sim.download()
    print(f"  rel {_rel(out, ref):.3e}", flush=True)
    assert np.isfinite(out).all()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("scheme,bc,nvz", [(O.FV, O.PERIODIC, 300), (O.FV, O.OPEN, 480),
                                           (O.FD, O.OPEN, 300)])
def test_general_rhs_matches_oracle(gpu, scheme, bc, nvz):
    po, pc = _pair(scheme, 32, nvz, bc=bc)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=46, amp=1e-3)
    ref = O.rhs(po, s)
    print(f"general rhs nvz={nvz} scheme={scheme} bc={bc}", flush=True)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        out = sim.rhs()
    assert _rel(out, ref) <= 1e-12


@pytest.mark.parametrize("zchunk", ["8", "1", "-2"])
def test_general_launch_shapes_agree(gpu, monkeypatch, zchunk):
    """Chunks of 8 rows, of 1 row (every FV range starts with the faces
    below it) and contiguous ranges over twice the resident wavefronts."""
    po, pc = _pair(O.FV, 64, 480, bc=O.OPEN)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=47, amp=1e-3)
    ref = O.step(po, s, 4)
    monkeypatch.setenv("COSENU_ZCHUNK", zchunk)
    print(f"general zchunk={zchunk}", flush=True)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(4)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_general_fast_agrees_with_strict(gpu, scheme):
    """Open boundary, mixed groups and padding over 60 steps of the
    flavour-unstable regime, against STRICT mode's per-cell loop."""
    po, pc = _pair(scheme, 256, 480, bc=O.OPEN)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=48, amp=1e-3)
    outs = {}
    for mode in (C.STRICT, C.FAST):
        pc.mode = mode
        print(f"general vs strict scheme={scheme} mode={mode}", flush=True)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(60)
            outs[mode] = sim.download()
    assert np.isfinite(outs[C.FAST]).all()
    assert _rel(outs[C.FAST], outs[C.STRICT]) <= 1e-10
