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
    Xtr, Ytr, Xte, Yte = X[:192], Y[:192], X[192:], Y[192:]
    cfg = S.TrainConfig(steps=800, batch=128, lr=2e-3, members=2)
    members = [S.train_member(Xtr, Ytr, cfg, seed=50 + k)[0] for k in range(cfg.members)]
    sur = S.Surrogate(members, bound=0.02)
    cal = sur.calibrate(Xte, Yte)
    base_err = float(np.abs(Ytr.mean(axis=0)[None, :] - Yte).mean())
    assert cal["heldout_mean_abs_err_all"] < 0.7 * base_err, (cal, base_err)
    # screened prediction: accepted samples from the network, the rest from the solver
    pred, accept = sur.screen(Xte)
    pred = pred.copy()
    if (~accept).any():
        idx = np.nonzero(~accept)[0]
        pred[idx], _ = S.solve([elns[192 + i] for i in idx], setup)
    err = np.abs(pred - Yte).mean()
    assert err <= 0.02 + 1e-6
    # a threshold below every spread sends every sample to the solver
    sur.tau = -1.0
    _, acc2 = sur.screen(Xte)
    assert not acc2.any()
    for m in members:
        m.close()
