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
    def __init__(self, d, seed=11):
        self.data_dir = str(d)
        self.seed = seed


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
