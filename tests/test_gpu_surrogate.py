"""The surrogate pipeline at small scale: solver-generated data, ensemble
training on the GPU kernels, held-out calibration, solver fallback, and the
gain over the method without a network.  (The full-size run is
`python -m cosenu_amd.train_surrogate`; its report is kept under profiles/.)"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from cosenu_amd import surrogate as S
    setup = S.SolverSetup(nz=256, z0=-150.0, z1=150.0, steps=600, sample_every=20)
    elns = S.sample_elns(224, seed=3)
    X = np.stack([S.features(e) for e in elns])
    Y, _ = S.solve(elns, setup)
    return S, setup, elns, X, Y


def test_targets_are_survival_probabilities(gpu, small):
    S, setup, elns, X, Y = small
    assert X.shape == (224, 67) and Y.shape == (224, 64)
    assert np.all(np.isfinite(Y)) and Y.min() > -0.05 and Y.max() < 1.05
    # flavour conversion happened in some runs and not in others
    assert Y.min() < 0.95 and np.ptp(Y.mean(axis=1)) > 0.02


def test_ensemble_beats_no_network_and_fallback_holds_the_bound(gpu, small):
    S, setup, elns, X, Y = small
    Xtr, Ytr = X[:160], Y[:160]
    Xca, Yca, Xte, Yte = X[160:192], Y[160:192], X[192:], Y[192:]
    cfg = S.TrainConfig(steps=800, batch=128, lr=2e-3, members=2)
    members = [S.train_member(Xtr, Ytr, cfg, seed=50 + k)[0] for k in range(cfg.members)]
    sur = S.Surrogate(members, bound=0.02)
    cal = sur.calibrate(Xca, Yca)
    # by construction no calibration run above the bound passes the per-run gate
    m = cal["run_gate_max_accepted_run_err"]
    assert m is None or m <= 0.02
    test = sur.evaluate(Xte, Yte)
    base_err = float(S.run_errors(Ytr.mean(axis=0)[None, :], Yte).mean())
    assert test["mean_err_network_only"] < 0.7 * base_err, (test, base_err)
    # screened prediction: accepted runs from the network, the rest from the solver
    for gate in ("run", "mean"):
        pred, accept = sur.screen(Xte, gate=gate)
        pred = pred.copy()
        if (~accept).any():
            idx = np.nonzero(~accept)[0]
            pred[idx], _ = S.solve([elns[192 + i] for i in idx], setup)
        err = S.run_errors(pred, Yte)
        assert err[~accept].max(initial=0.0) == 0.0  # the solver's own values
        assert err.mean() == pytest.approx(test[f"{gate}_gate"]["mean_err_screened"], abs=1e-6)
        assert err.mean() <= test["mean_err_network_only"] + 1e-9
        if gate == "run":
            assert err.mean() <= 0.02
    # a threshold below every spread sends every run to the solver
    sur.tau = -1.0
    _, acc2 = sur.screen(Xte)
    assert not acc2.any()
    for m in members:
        m.close()


def test_ensemble_file_round_trip(gpu, small, tmp_path):
    S, setup, elns, X, Y = small
    cfg = S.TrainConfig(steps=50, batch=64, members=2)
    members = [S.train_member(X[:64], Y[:64], cfg, seed=60 + k)[0] for k in range(2)]
    sur = S.Surrogate(members, bound=0.02, meta={"sweep_seed": 5})
    sur.calibrate(X[64:128], Y[64:128])
    path = str(tmp_path / "ens.npz")
    sur.save(path)
    back = S.Surrogate.load(path)
    m1, s1 = sur.predict(X[128:])
    m2, s2 = back.predict(X[128:])
    assert np.array_equal(m1, m2) and np.array_equal(s1, s2)
    assert back.tau == sur.tau and back.meta == {"sweep_seed": 5}
    for m in members + back.members:
        m.close()
