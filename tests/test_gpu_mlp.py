"""The surrogate network's GPU kernels (cosenu_mlp_*) against the CPU FP64
restatement (oracle/mlp_oracle.py).

With the oracle rounding to bfloat16 exactly where the GPU does, the only
remaining difference is FP32 (GPU) versus FP64 (CPU) accumulation, so the
bars are tight: products and forward pass 1e-5 relative for a single layer,
and normwise 5e-3 for the 4-layer forward pass and its gradient, where a
pre-activation that lands within FP32 rounding of a bfloat16 rounding
boundary can move one activation by one bfloat16 step.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _S():
    from cosenu_amd import surrogate
    return surrogate


def _norm(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("M_,K,N", [(16, 16, 16), (37, 67, 50), (100, 256, 64), (5, 3, 33)])
def test_single_layer_matches_bf16_product(gpu, M_, K, N):
    """One linear layer: checks the MFMA operand/accumulator lane layout and
    the bounds handling of ragged tiles."""
    S = _S()
    rng = np.random.default_rng(M_ * 1000 + K)
    net = S.MLP((K, N), max_batch=128)
    flat = rng.standard_normal(K * N + N).astype(np.float32)
    net.set_params(flat)
    X = rng.standard_normal((M_, K)).astype(np.float32)
    Y = net.forward(X)
    ref, _ = M.forward(flat, [K, N], X, bf16=True)
    assert np.max(np.abs(Y - ref)) <= 1e-5 * np.max(np.abs(ref))
    net.close()


def test_forward_matches_oracle(gpu):
    S = _S()
    rng = np.random.default_rng(1)
    net = S.MLP(S.DIMS, max_batch=512)
    flat = S.init_params(S.DIMS, seed=2)
    flat[-64:] = 0.3
    net.set_params(flat)
    X = rng.standard_normal((300, 67)).astype(np.float32)
    Y = net.forward(X)
    ref, _ = M.forward(flat, list(S.DIMS), X, bf16=True)
    assert _norm(Y, ref) <= 5e-3
    ref64, _ = M.forward(flat, list(S.DIMS), X, bf16=False)
    assert _norm(Y, ref64) <= 3e-2  # what BF16 storage costs against pure FP64
    assert np.array_equal(net.get_params(), flat)
    net.close()


def test_gradients_match_oracle(gpu):
    S = _S()
    rng = np.random.default_rng(3)
    net = S.MLP(S.DIMS, max_batch=512)
    flat = S.init_params(S.DIMS, seed=4)
    net.set_params(flat)
    X = rng.standard_normal((200, 67)).astype(np.float32)
    T = rng.random((200, 64)).astype(np.float32)
    loss, g = net.gradients(X, T)
    ref_loss, ref_g = M.loss_and_grad(flat, list(S.DIMS), X, T, bf16=True)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    assert _norm(g, ref_g) <= 5e-3
    # layer by layer, so that one wrong block cannot hide behind the others
    o = 0
    for k, n in zip(S.DIMS[:-1], S.DIMS[1:]):
        for size in (k * n, n):
            assert _norm(g[o:o + size], ref_g[o:o + size]) <= 1e-2
            o += size
    _, g64 = M.loss_and_grad(flat, list(S.DIMS), X, T, bf16=False)
    assert _norm(g, g64) <= 5e-2
    net.close()


def test_adamw_step_matches_oracle(gpu):
    S = _S()
    rng = np.random.default_rng(5)
    dims = (67, 64, 64)
    net = S.MLP(dims, max_batch=64)
    flat = S.init_params(dims, seed=6)
    net.set_params(flat)
    X = rng.standard_normal((50, 67)).astype(np.float32)
    T = rng.random((50, 64)).astype(np.float32)
    _, g = net.gradients(X, T)
    m = v = np.zeros(flat.size)
    ref = flat.astype(np.float64)
    for t in (1, 2):  # the second step uses the moments of the first
        _, g = net.gradients(X, T)
        net.train_step(X, T, 1e-3, weight_decay=0.01)
        ref, m, v = M.adamw_step(ref, g.astype(np.float64), m, v, t, list(dims), 1e-3, wd=0.01)
        assert np.max(np.abs(net.get_params() - ref)) <= 1e-6
    net.close()


def test_training_fits_a_smooth_map(gpu):
    """Loss falls by orders of magnitude on a smooth 67 -> 64 regression."""
    S = _S()
    rng = np.random.default_rng(7)
    X = rng.uniform(-1, 1, (1024, 67)).astype(np.float32)
    A = rng.standard_normal((67, 64)) / 8
    Y = (0.5 + 0.4 * np.tanh(X @ A)).astype(np.float32)
    cfg = S.TrainConfig(steps=600, batch=256, lr=3e-3, lr_min=1e-4, weight_decay=0.0)
    net, losses = S.train_member(X, Y, cfg, seed=1)
    assert losses[-1][1] < 0.05 * losses[0][1]
    net.close()
