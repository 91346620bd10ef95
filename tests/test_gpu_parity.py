"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

STRICT mode must be bit-identical to the oracle (which is itself bit-identical
to the reference build, tests/test_oracle_ref.py).  FAST mode evaluates the
same mathematics with a different operation order (moment-form collective
integral, one-sided WENO with one division, fused multiply-adds); it must
agree normwise: ||gpu - ref||_inf / ||ref||_inf <= 1e-12 for one RHS of a
smooth state (1e-9 for a state with 1e-4 noise, where the reference's own
arithmetic is ill-conditioned) and <= 1e-10 after RK4 steps.
"""
import math

import numpy as np
import pytest

import oracle as O
import cosenu_amd as C
from cosenu_amd.initial import seeded_ic

pytestmark = pytest.mark.gpu

VARIANTS = [
    (O.FV, O.PERIODIC, O.VAC | O.COLL),
    (O.FV, O.PERIODIC, O.VAC),
    (O.FV, O.PERIODIC, O.COLL),
    (O.FV, O.PERIODIC, 0),
    (O.FV, O.PERIODIC, O.VAC | O.MAT | O.COLL),
    (O.FV, O.PERIODIC, O.VAC | O.COLL | O.ADVEC_OFF),
    (O.FV, O.OPEN, O.VAC | O.COLL),
    (O.FD, O.PERIODIC, O.VAC | O.COLL),
    (O.FD, O.PERIODIC, O.VAC | O.MAT | O.COLL),
    (O.FD, O.PERIODIC, 0),
    (O.FD, O.OPEN, O.VAC | O.COLL),
    (O.FD, O.PERIODIC, O.VAC | O.COLL | O.KO5),
]


def _pair(scheme, bc, flags, nz, nvz, mode, omega=0.3, theta=37.0):
    kw = dict(nz=nz, nvz=nvz, z0=-nz / 2.0, z1=nz / 2.0, scheme=scheme, bc=bc, flags=flags,
              omega=omega, theta=theta, pmo=1.0, mu=1.0)
    return O.RunParams(**kw), C.RunParams(mode=mode, **kw)


def _state(po, seed, amp=1e-3):
    G, bG, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=seed, amp=amp)
    rng = np.random.default_rng(seed)
    return G, bG, s + 1e-4 * rng.standard_normal(s.shape)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("nvz", [12, 64])
@pytest.mark.parametrize("scheme,bc,flags", VARIANTS)
def test_strict_rhs_bitwise(gpu, scheme, bc, flags, nvz):
    po, pc = _pair(scheme, bc, flags, 48, nvz, C.STRICT)
    _, _, s = _state(po, 11)
    ref = O.rhs(po, s)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        out = sim.rhs()
    assert np.array_equal(out, ref), _rel(out, ref)


@pytest.mark.parametrize("scheme,bc,flags", VARIANTS)
def test_strict_rk4_bitwise(gpu, scheme, bc, flags):
    po, pc = _pair(scheme, bc, flags, 64, 20, C.STRICT)
    _, _, s = _state(po, 12)
    ref = O.step(po, s, 10)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(10)
        out = sim.download()
    assert np.array_equal(out, ref), _rel(out, ref)


def test_strict_rk4_bitwise_long_graph_path(gpu):
    """40 steps go through the captured hipGraph (16-step graphs) and the
    eager tail; still bit-identical."""
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 128, 64, C.STRICT, omega=0.0, theta=1e-6)
    _, _, s = _state(po, 13, amp=1e-6)
    ref = O.step(po, s, 40)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(40)
        out = sim.download()
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("nvz", [16, 64, 128, 256])
@pytest.mark.parametrize("scheme,flags", [(O.FV, O.VAC | O.COLL), (O.FV, O.VAC | O.MAT | O.COLL),
                                          (O.FV, O.COLL | O.ADVEC_OFF), (O.FD, O.VAC | O.COLL),
                                          (O.FD, O.VAC | O.MAT | O.COLL)])
def test_fast_rhs_normwise(gpu, scheme, flags, nvz):
    """Smooth physical state: FAST agrees with the reference arithmetic to
    1e-12 normwise.  Noisy state (1e-4 noise on O(1) values): the reference's
    expanded smoothness indicators lose ~8 digits to cancellation (its own
    error against a long-double evaluation reaches 4e-11 per flux,
    tests/test_oracle_math.py::test_reference_weno_conditioning), so the bar
    there is 1e-9."""
    po, pc = _pair(scheme, O.PERIODIC, flags, 64, nvz, C.FAST)
    G, bG, smooth = seeded_ic(po, 0.6, 0.5, 0.9, seed=14, amp=1e-3)
    _, _, noisy = _state(po, 14)
    with C.Simulation(pc) as sim:
        for s, tol in ((smooth, 1e-12), (noisy, 1e-9)):
            ref = O.rhs(po, s)
            sim.upload(s)
            out = sim.rhs()
            assert _rel(out, ref) <= tol


@pytest.mark.parametrize("nvz", [16, 40, 64, 100, 128, 256])
@pytest.mark.parametrize("scheme", [O.FV, O.FD])
def test_fast_rk4_normwise(gpu, scheme, nvz):
    """nvz not a multiple of 64 leaves padding bins in the device layout;
    uploads stage host data through the stage buffers, so those bins hold
    stale values that must never reach the moments (regression test)."""
    po, pc = _pair(scheme, O.PERIODIC, O.VAC | O.COLL, 128, nvz, C.FAST)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=15, amp=1e-3)
    ref = O.step(po, s, 30)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.step(30)
        out = sim.download()
    assert _rel(out, ref) <= 1e-10


def test_fast_deterministic_and_graph_equals_eager(gpu, monkeypatch):
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 256, 128, C.FAST)
    _, _, s = _state(po, 16)
    outs = []
    for gsteps in ("16", "0", "16"):
        monkeypatch.setenv("COSENU_GRAPH_STEPS", gsteps)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(32)
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("scheme,variant", [(O.FV, "4"), (O.FD, "5"), (O.FV, "3")])
def test_prepared_graphs_equal_eager_steps(gpu, monkeypatch, scheme, variant):
    """cosenu_prepare captures the step graph for both state buffers the two-
    and four-stage launches alternate; after an odd number of steps the
    graph of the other buffer is replayed.  Results equal steps launched one
    by one, bit for bit."""
    monkeypatch.setenv("COSENU_FAST_VARIANT", variant)
    po, pc = _pair(scheme, O.PERIODIC, O.VAC | O.COLL, 256, 64, C.FAST)
    _, _, s = _state(po, 23)
    outs = []
    for gsteps in ("16", "0"):
        monkeypatch.setenv("COSENU_GRAPH_STEPS", gsteps)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(variant)
            sim.upload(s)
            sim.prepare()
            sim.step(3)   # odd: the state now starts in the other buffer
            sim.step(16)  # one graph replay from that buffer
            sim.step(17)  # one replay from the first buffer, one step alone
            outs.append(sim.download())
    assert np.array_equal(outs[0], outs[1])


def test_layout_round_trips(gpu):
    _, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 72, 40, C.STRICT)
    rng = np.random.default_rng(3)
    s = rng.standard_normal((8, pc.nvz, pc.nz))
    with C.Simulation(pc) as sim:
        sim.upload(s)
        assert np.array_equal(sim.download(), s)
        inter = sim.download_interleaved()
        assert np.array_equal(inter, np.moveaxis(s, 0, -1))
        sim.upload_interleaved(np.ascontiguousarray(np.moveaxis(s[::-1], 0, -1)), C.STATE0)
        assert np.array_equal(sim.download(C.STATE0), s[::-1])


def test_fieldvar_layout_with_ghost_cells(gpu):
    """The reference's FieldVar rows carry gz ghost cells on each side
    (nuosc.hpp:145-148); upload/download address the interior only."""
    import ctypes
    _, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 40, 24, C.STRICT)
    gz, R = 4, 40 + 8
    rng = np.random.default_rng(4)
    planes = [np.ascontiguousarray(rng.standard_normal((pc.nvz, R))) for _ in range(8)]
    with C.Simulation(pc) as sim:
        arr = C._planes_t(*[p.ctypes.data_as(C._dp) for p in planes])
        assert C.lib().cosenu_upload_planes(sim._h, C.STATE, arr, R, gz) == 0
        out = sim.download()
        for f in range(8):
            assert np.array_equal(out[f], planes[f][:, gz:gz + 40])
        back = [np.full((pc.nvz, R), 7.0) for _ in range(8)]
        arr2 = C._planes_t(*[p.ctypes.data_as(C._dp) for p in back])
        assert C.lib().cosenu_download_planes(sim._h, C.STATE, arr2, R, gz) == 0
        for f in range(8):
            assert np.array_equal(back[f][:, gz:gz + 40], planes[f][:, gz:gz + 40])
            assert np.all(back[f][:, :gz] == 7.0) and np.all(back[f][:, gz + 40:] == 7.0)


@pytest.mark.parametrize("mode", [C.STRICT, C.FAST])
def test_diagnostics_match_oracle(gpu, mode):
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 128, 64, mode, omega=0.0, theta=1e-6)
    G, bG, s0 = _state(po, 17, amp=1e-2)
    s = O.step(po, s0, 20)
    ref = O.diagnostics(po, G, bG, s0, s)
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(s0, C.STATE0)
        sim.upload(s)
        d = sim.diagnostics()
        pee, pbee = sim.survival_profile()
    for k, v in ref.items():
        assert d[k] == pytest.approx(v, rel=1e-12, abs=1e-14), k
    assert np.allclose(pee, s[0].sum(1) / s0[0].sum(1), rtol=1e-13)
    assert np.allclose(pbee, s[4].sum(1) / s0[4].sum(1), rtol=1e-13)


def test_nonfinite_state_is_flagged(gpu):
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 64, 64, C.FAST)
    G, bG, s0 = _state(po, 18)
    s = s0.copy()
    s[3, 5, 7] = np.nan
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(s0, C.STATE0)
        sim.upload(s)
        d = sim.diagnostics()
    assert d["nonfinite"] == 1.0 and np.isnan(d["dP_max"])


def test_snapshot_pee_async(gpu):
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 256, 64, C.FAST)
    G, bG, s0 = _state(po, 19)
    with C.Simulation(pc) as sim:
        sim.set_profile(G, bG)
        sim.upload(s0, C.STATE0)
        sim.upload(s0)
        sim.step(3)
        buf = C.PinnedBuffer(pc.nvz * pc.nz)
        try:
            assert C.lib().cosenu_snapshot_pee_async(sim._h, buf.ptr) == 0
            sim.step(2)  # keeps computing while the copy runs
            assert C.lib().cosenu_snapshot_wait(sim._h) == 0
            snap = buf.array.reshape(pc.nvz, pc.nz).copy()
        finally:
            buf.free()
        # the snapshot is the state after 3 steps
        sim2 = C.Simulation(pc)
        sim2.upload(s0)
        sim2.step(3)
        s3 = sim2.download()
        sim2.close()
    assert np.array_equal(snap, s3[0] / s0[0])


@pytest.mark.parametrize("variant", ["2", "4"])
def test_timing_reports_every_launch(gpu, monkeypatch, variant):
    """One timing slot per launch of a step: four one-stage launches, or two
    two-stage launches."""
    monkeypatch.setenv("COSENU_FAST_VARIANT", variant)
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 512, 128, C.FAST)
    _, _, s = _state(po, 20)
    with C.Simulation(pc) as sim:
        sim.upload(s)
        sim.set_timing(True)
        sim.step(5)
        ms, cnt = sim.timing()
        names = [sim.stage_kernel_name(k) for k in range(4)]
    n = 4 if variant == "2" else 2
    assert cnt == [5] * n + [0] * (4 - n)
    assert all(m > 0 for m in ms[:n]) and all(names[:n]) and not any(names[n:])


def test_errors_are_reported_not_fatal(gpu):
    with pytest.raises(C.CosenuError, match="nvz"):
        C.Simulation(C.RunParams(nz=64, nvz=0))
    with pytest.raises(C.CosenuError, match="nz"):
        C.Simulation(C.RunParams(nz=4, nvz=512, mode=C.FAST, bc=C.OPEN))
    with C.Simulation(C.RunParams(nz=64, nvz=16, mode=C.STRICT)) as sim:
        with pytest.raises(C.CosenuError, match="profile"):
            sim.diagnostics()
        # host buffers of the wrong shape never reach the library
        with pytest.raises(ValueError):
            sim.upload(np.zeros((8, 16, 63)))
        with pytest.raises(ValueError):
            sim.set_profile_v(np.ones(15), np.ones(16))


def test_separable_initial_state_equals_host_state(gpu):
    """cosenu_init_separable / cosenu_set_profile_v build on the device what
    the host computes (one rounded product per value)."""
    from cosenu_amd.initial import seeded_factors
    po, pc = _pair(O.FV, O.PERIODIC, O.VAC | O.COLL, 96, 40, C.FAST)
    a, b, G1, bG1 = seeded_factors(pc, 0.6, 0.5, 0.9, seed=21, amp=1e-3, gaussian_seed=True)
    G, bG, s0 = seeded_ic(pc, 0.6, 0.5, 0.9, seed=21, amp=1e-3, gaussian_seed=True)
    with C.Simulation(pc) as sim:
        sim.init_separable(a, b, C.STATE)
        sim.init_separable(a, b, C.STATE0)
        assert np.array_equal(sim.download(), s0)
        assert np.array_equal(sim.download(C.STATE0), s0)
        sim.set_profile_v(G1, bG1)
        d1 = sim.diagnostics()
        sim.set_profile(G, bG)
        d2 = sim.diagnostics()
    assert d1 == d2


@pytest.mark.parametrize("scheme,bc,flags", [(O.FV, O.OPEN, O.VAC | O.COLL),
                                             (O.FD, O.OPEN, O.VAC | O.COLL),
                                             (O.FD, O.PERIODIC, O.VAC | O.COLL | O.KO5),
                                             (O.FV, O.OPEN, O.VAC | O.MAT | O.COLL)])
@pytest.mark.parametrize("nvz", [24, 128])
def test_fast_open_boundary_and_ko5(gpu, scheme, bc, flags, nvz):
    """FAST mode with the open boundary of nuosc.hpp:247-277 (ghost rows and
    flux ghosts) and the 5th-order Kreiss-Oliger term of rhs_fd.hpp:129-133."""
    po, pc = _pair(scheme, bc, flags, 96, nvz, C.FAST)
    _, _, s = seeded_ic(po, 0.6, 0.5, 0.9, seed=22, amp=1e-3)
    # a z-dependent background so the boundary rules matter
    s = s * (1.0 + 0.05 * np.sin(np.linspace(0, 3.0, po.nz)))[None, None, :]
    with C.Simulation(pc) as sim:
        sim.upload(s)
        assert _rel(sim.rhs(), O.rhs(po, s)) <= 1e-12
        sim.upload(s)
        sim.step(30)
        out = sim.download()
    assert _rel(out, O.step(po, s, 30)) <= 1e-10


@pytest.mark.parametrize("variant", ["2", "4"])
def test_fast_flavor_growth_rate_matches_linear_theory(gpu, monkeypatch, variant):
    """The FAST kernels reproduce the linear growth and turning rate of the
    homogeneous fast-flavour instability (the same check as the oracle's,
    tests/test_oracle_math.py::test_fast_flavor_growth_rate_matches_linear_theory)."""
    from test_oracle_math import ffi_growth_check
    monkeypatch.setenv("COSENU_FAST_VARIANT", variant)

    def gpu_step(p, s, n):
        pc = C.RunParams(nz=p.nz, nvz=p.nvz, z0=p.z0, z1=p.z1, CFL=p.CFL, flags=p.flags,
                         mu=p.mu, mode=C.FAST)
        with C.Simulation(pc) as sim:
            assert sim.info()["fast_variant"] == int(variant)
            sim.upload(s)
            sim.step(n)
            return sim.download()

    rate, lam, resid, t = ffi_growth_check(gpu_step)
    assert abs(rate.real - lam.real) <= 1e-5 * lam.real
    turn = (lam.imag - rate.imag) * t / (2 * math.pi)
    assert abs(turn - round(turn)) < 1e-6
    assert resid < 1e-12


@pytest.mark.parametrize("variant", ["2", "4"])
def test_polarisation_length_conserved_up_to_rk4_truncation(gpu, monkeypatch, variant):
    """As the oracle's test (tests/test_oracle_math.py): |P| per cell changes
    only by the RK4 truncation error under the non-advective terms."""
    from test_oracle_math import plength_drift
    monkeypatch.setenv("COSENU_FAST_VARIANT", variant)

    def gpu_step(p, s, n):
        pc = C.RunParams(nz=p.nz, nvz=p.nvz, z0=p.z0, z1=p.z1, CFL=p.CFL, flags=p.flags,
                         omega=p.omega, theta=p.theta, mu=p.mu, mode=C.FAST)
        with C.Simulation(pc) as sim:
            sim.upload(s)
            sim.step(n)
            return sim.download()

    e1, e2 = plength_drift(gpu_step, 0.02), plength_drift(gpu_step, 0.01)
    assert e2 < 1e-10
    assert 12.0 < e1 / e2 < 20.0, (e1, e2)
