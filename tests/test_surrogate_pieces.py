"""Host logic of the surrogate driver's dataset pieces (no GPU): runs solved
in pieces and loaded back give the same dataset as one solve, pieces of
another grid or seed are ignored, and missing runs are solved."""
import json

import numpy as np

import cosenu_amd.surrogate as S
import cosenu_amd.train_surrogate as T


def _fake_solve(calls):
    def solve(elns, setup, log=None):
        calls.append(len(elns))
        # a deterministic function of the run and the grid
        Y = np.array([[e.sig_nu + e.alpha * k + setup.nz * 1e-6 + e.seed * 1e-9 for k in range(S.NV)]
                      for e in elns], dtype=np.float32)
        return Y, 0.5 * len(elns)
    return solve


class _Args:
    def __init__(self, d, seed=11, runs=12, dataset=""):
        self.data_dir = str(d)
        self.seed = seed
        self.runs = runs
        self.dataset = dataset


def test_pieces_reassemble_the_dataset(tmp_path, monkeypatch):
    calls = []
    monkeypatch.setattr(S, "solve", _fake_solve(calls))
    setup = S.SolverSetup(nz=64, steps=10)
    args = _Args(tmp_path)
    T.save_piece(args, setup, 0, 5)
    T.save_piece(args, setup, 5, 9)
    assert calls == [5, 4]
    elns = S.sample_elns(12, args.seed)  # a larger sample: runs 9..11 are missing
    Y, t, pieces = T.dataset(args, setup, elns)
    assert calls == [5, 4, 3]
    want, _ = _fake_solve([])(elns, setup)
    assert np.array_equal(Y, want)
    assert [p[:2] for p in pieces] == [[0, 5], [5, 9], [9, 12]]
    assert t == 0.5 * 12


def test_pieces_of_another_grid_or_seed_are_ignored(tmp_path, monkeypatch):
    calls = []
    monkeypatch.setattr(S, "solve", _fake_solve(calls))
    T.save_piece(_Args(tmp_path), S.SolverSetup(nz=128, steps=10), 0, 4)
    T.save_piece(_Args(tmp_path, seed=12), S.SolverSetup(nz=64, steps=10), 0, 4)
    setup = S.SolverSetup(nz=64, steps=10)
    elns = S.sample_elns(4, 11)
    Y, _, pieces = T.dataset(_Args(tmp_path), setup, elns)
    assert calls[-1] == 4 and pieces == [[0, 4, "solved here"]]
    want, _ = _fake_solve([])(elns, setup)
    assert np.array_equal(Y, want)


def test_sample_prefix_is_stable():
    a = S.sample_elns(7, 3)
    b = S.sample_elns(20, 3)[:7]
    assert a == b


def test_collected_dataset_file_round_trip(tmp_path, monkeypatch):
    """--collect gathers the pieces (ELN parameters and targets) into one file
    that --dataset reads back; a file of another seed is refused."""
    calls = []
    monkeypatch.setattr(S, "solve", _fake_solve(calls))
    setup = S.SolverSetup(nz=64, steps=10)
    args = _Args(tmp_path / "pieces", runs=9)
    T.save_piece(args, setup, 0, 5)
    T.save_piece(args, setup, 5, 9)
    path = str(tmp_path / "all.npz")
    T.collect(args, setup, path)
    elns = S.sample_elns(9, args.seed)
    want, _ = _fake_solve([])(elns, setup)
    Y, have, t = T.load_dataset_file(path, args, setup, elns)
    assert have.all() and np.array_equal(Y, want) and t == 0.5 * 9
    n = len(calls)
    Y2, _, pieces = T.dataset(_Args(tmp_path / "none", runs=9, dataset=path), setup, elns)
    assert len(calls) == n and np.array_equal(Y2, want) and pieces == [[0, 9, "all.npz"]]
    import pytest
    with pytest.raises(SystemExit):
        T.load_dataset_file(path, _Args(tmp_path, seed=12, runs=9), setup, S.sample_elns(9, 12))


class _Member:
    """Stands in for an MLP: fixed outputs per input row."""

    def __init__(self, out):
        self.out = np.asarray(out, dtype=np.float32)

    def forward(self, X):
        return self.out[np.asarray(X[:, 0], dtype=int)]


def test_gate_rejects_every_calibration_run_above_the_bound():
    """tau sits just below the smallest ensemble spread of a calibration run
    whose own error exceeds the bound; evaluate() counts accepted runs above
    the bound and the recall of such runs on further data."""
    nv = 4
    truth = np.full((6, nv), 0.5, dtype=np.float32)
    # member outputs: run k has both members at 0.5 +- d_k (spread d_k) and
    # an offset e_k of the mean (error e_k)
    d = np.array([0.001, 0.002, 0.003, 0.010, 0.004, 0.020])
    e = np.array([0.005, 0.010, 0.030, 0.040, 0.015, 0.001])
    a = truth + (e + d)[:, None]
    b = truth + (e - d)[:, None]
    sur = S.Surrogate([_Member(a), _Member(b)], bound=0.02)
    X = np.arange(6, dtype=np.float32)[:, None]
    cal = sur.calibrate(X, truth)
    # runs 2 and 3 exceed the bound; the smaller of their spreads is 0.003
    assert sur.tau < 0.003 and sur.tau > 0.0029999
    assert cal["above_bound"] == 2 and cal["run_gate_max_accepted_run_err"] <= 0.02
    assert cal["run_gate_accept_rate"] == 2 / 6
    ev = sur.evaluate(X, truth)
    g = ev["run_gate"]
    assert g["accepted_above_bound"] == 0 and g["recall_of_above_bound"] == 1.0
    assert np.isclose(g["mean_err_screened"], (0.005 + 0.010) / 6)
    # sweep-mean gate: runs accepted in order of spread while the screened
    # mean (rejected runs exact) stays within the bound -- here all of them,
    # 0.101 / 6 = 0.0168
    assert np.isclose(sur.tau_mean, 0.020) and cal["mean_gate_accept_rate"] == 1.0
    assert np.isclose(ev["mean_gate"]["mean_err_screened"], 0.101 / 6)
    _, acc = sur.screen(X, gate="mean")
    assert acc.all()
    # a tighter bound: the three least-spread runs (0.045 / 6 <= 0.009)
    sur3 = S.Surrogate([_Member(a), _Member(b)], bound=0.009)
    sur3.calibrate(X, truth)
    assert np.isclose(sur3.tau_mean, 0.003)
    # no run above the bound: every calibration spread is accepted
    sur2 = S.Surrogate([_Member(truth + 0.001), _Member(truth - 0.001)], bound=0.02)
    cal2 = sur2.calibrate(X, truth)
    assert cal2["run_gate_accept_rate"] == 1.0 and cal2["above_bound"] == 0
