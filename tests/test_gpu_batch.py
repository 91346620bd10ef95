"""Batched runs (cosenu_batch_*) and the autotuner.

A batch of runs with uneven nz, nvz and physics must leave every run
bit-identical to the same run advanced alone in its own context with the same
FAST kernel (variant 5, the default: FD groups four stages per launch, the
others two; variant 4: two stages per launch everywhere,
COSENU_BATCH_VARIANT=4; variant 2: one stage per launch,
COSENU_BATCH_VARIANT=2), and its diagnostics equal too.  Runs of one batch
are grouped by (nvz rounded to 64, scheme).
"""
import os

import numpy as np
import pytest

import cosenu_amd as C
import oracle as O
from cosenu_amd.initial import seeded_ic

pytestmark = pytest.mark.gpu


def _runs():
    rng = np.random.default_rng(42)
    specs = [(256, 64, C.FV), (448, 128, C.FV), (128, 64, C.FV), (320, 128, C.FV),
             (192, 64, C.FD), (512, 256, C.FV), (64, 128, C.FV)]
    runs, ics = [], []
    for k, (nz, nvz, scheme) in enumerate(specs):
        p = C.RunParams(nz=nz, nvz=nvz, z0=-nz / 4.0, z1=nz / 4.0, scheme=scheme,
                        flags=C.VAC | C.COLL, omega=float(rng.uniform(0, 0.5)),
                        theta=float(rng.uniform(1, 40)), mu=float(rng.uniform(0.5, 1.5)),
                        mode=C.FAST)
        ics.append(seeded_ic(p, rng.uniform(0.4, 0.8), rng.uniform(0.4, 0.8),
                             rng.uniform(0.7, 1.1), seed=100 + k, amp=1e-3))
        runs.append(p)
    return runs, ics


@pytest.fixture(params=["2", "4", "5"])
def variant(request, monkeypatch):
    monkeypatch.setenv("COSENU_FAST_VARIANT", request.param)
    monkeypatch.setenv("COSENU_BATCH_VARIANT", request.param)
    return int(request.param)


@pytest.mark.parametrize("chunk", ["", "48"])
@pytest.mark.parametrize("nsteps", [20, 37])
def test_batch_runs_bit_identical_to_solo_runs(gpu, variant, nsteps, chunk, monkeypatch):
    # 20 = one 16-step graph replay + 4 eager steps; 37 (odd) leaves the
    # two-stage batches' states in their second buffers.  Chunks of 48 rows
    # divide none of the runs' nz but 192: each run's last item is shorter.
    if chunk:
        monkeypatch.setenv("COSENU_BATCH_ZCHUNK", chunk)
    runs, ics = _runs()
    with C.Batch(runs) as b:
        info = b.info()
        assert info["nruns"] == len(runs) and info["ngroups"] == 4
        if chunk:
            assert info["z_chunk"] == int(chunk)
        assert info["fast_variant"] == variant
        assert info["cells"] == sum(p.nz * p.nvz for p in runs)
        for r, (G, bG, s0) in enumerate(ics):
            b.set_profile(r, G, bG)
            b.upload(r, s0, C.STATE0)
            b.upload(r, s0)
        b.step(nsteps)
        b.sync()
        outs = [b.download(r) for r in range(len(runs))]
        diags = [b.diagnostics(r) for r in range(len(runs))]
    for r, (p, (G, bG, s0)) in enumerate(zip(runs, ics)):
        with C.Simulation(p) as sim:
            sim.set_profile(G, bG)
            sim.upload(s0, C.STATE0)
            sim.upload(s0)
            sim.step(nsteps)
            solo = sim.download()
            d = sim.diagnostics()
        assert np.array_equal(outs[r], solo), r
        assert diags[r] == d, r
        assert not np.array_equal(solo, s0)


def test_batch_matches_oracle(gpu, variant):
    runs, ics = _runs()
    runs, ics = runs[:3], ics[:3]
    with C.Batch(runs) as b:
        for r, (G, bG, s0) in enumerate(ics):
            b.upload(r, s0)
        b.step(10)
        outs = [b.download(r) for r in range(len(runs))]
    for p, (G, bG, s0), out in zip(runs, ics, outs):
        po = O.RunParams(nz=p.nz, nvz=p.nvz, z0=p.z0, z1=p.z1, scheme=p.scheme, flags=p.flags,
                         omega=p.omega, theta=p.theta, mu=p.mu)
        ref = O.step(po, s0, 10)
        assert np.max(np.abs(out - ref)) / np.max(np.abs(ref)) <= 1e-10


def test_batch_rejects_bad_runs(gpu):
    with pytest.raises(C.CosenuError, match="run 1"):
        C.Batch([C.RunParams(nz=64, nvz=64), C.RunParams(nz=64, nvz=512)])
    with pytest.raises(C.CosenuError):
        C.Batch([])


def test_batch_autotune_records_and_applies(gpu, tmp_path, monkeypatch):
    table = tmp_path / "tune.txt"
    monkeypatch.setenv("COSENU_TUNE_FILE", str(table))
    monkeypatch.delenv("COSENU_BATCH_ZCHUNK", raising=False)
    monkeypatch.delenv("COSENU_BATCH_VARIANT", raising=False)
    runs, ics = _runs()
    L, ms = C.batch_autotune(runs)
    assert L in (32, 64, 128, 256, 512) and ms > 0
    rows = [l.split() for l in table.read_text().splitlines() if not l.startswith("#")]
    groups = {(r[0], r[1]) for r in rows if r[2] == "-1"}
    assert groups == {("64", "0"), ("128", "0"), ("64", "1"), ("256", "0")}
    assert all(r[4] == str(L) for r in rows if r[2] == "-1")
    with C.Batch(runs) as b:
        # the first group (nvz 64, FV: nz 256 and 128) caps the chunk at its largest nz
        assert b.info()["z_chunk"] == min(L, 256)
        for r, (G, bG, s0) in enumerate(ics):
            b.upload(r, s0)
        b.step(6)
        out = b.download(0)
    p = runs[0]
    G, bG, s0 = ics[0]
    monkeypatch.setenv("COSENU_FAST_VARIANT", "4")  # the batch's kernel
    with C.Simulation(p) as sim:
        sim.upload(s0)
        sim.step(6)
        assert np.array_equal(out, sim.download())


def test_autotune_records_and_applies(gpu, tmp_path, monkeypatch):
    table = tmp_path / "tune.txt"
    monkeypatch.setenv("COSENU_TUNE_FILE", str(table))
    monkeypatch.delenv("COSENU_FAST_VARIANT", raising=False)
    monkeypatch.delenv("COSENU_ZCHUNK", raising=False)
    p = C.RunParams(nz=4096, nvz=128, mode=C.FAST)
    v, L, ms = C.autotune(p)
    assert v in (1, 2, 3, 4) and ms > 0
    assert (L > 0 and 4096 % L == 0) or (v >= 2 and L in (-1, -2))
    rows = [l.split() for l in table.read_text().splitlines() if not l.startswith("#")]
    # (the table is process-wide: entries other tests tuned are written too)
    assert [["128", "0", "4096", str(v), str(L)]] == [r[:5] for r in rows
                                                      if r[1] == "0" and r[2] == "4096"]
    with C.Simulation(p) as sim:
        info = sim.info()
    assert info["fast_variant"] == v
    if L > 0:
        assert info["z_chunk"] == L and info["range_teams"] == 0
    else:
        assert info["range_teams"] > 0


@pytest.mark.parametrize("variant", [2, 3])
@pytest.mark.parametrize("mult", [1, 2])
def test_range_launch_equals_chunked_launch(gpu, monkeypatch, mult, variant):
    """The contiguous-range launch computes the same cells with the same
    arithmetic as the chunked one: bit-identical states."""
    monkeypatch.setenv("COSENU_FAST_VARIANT", str(variant))
    p = C.RunParams(nz=6000, nvz=128, z0=-300, z1=300, mode=C.FAST)
    _, _, s0 = seeded_ic(p, 0.6, 0.5, 0.9, seed=9, amp=1e-3)
    outs = []
    for chunk in ("40", str(-mult)):
        monkeypatch.setenv("COSENU_ZCHUNK", chunk)
        with C.Simulation(p) as sim:
            info = sim.info()
            sim.upload(s0)
            sim.step(20)
            outs.append(sim.download())
        assert (info["range_teams"] > 0) == (chunk.startswith("-"))
        assert info["fast_variant"] == variant
    assert np.array_equal(outs[0], outs[1])
