"""Pins the CPU restatement of the surrogate network (oracle/mlp_oracle.py):
its analytic gradient against central finite differences of its own loss,
the GELU derivative against differences of GELU, the bfloat16 rounding
against hand-computed bit patterns, and one AdamW step against the closed
form of the first step (update = -lr * sign(g) for bias-corrected moments)."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_bf16_rounding_bit_patterns():
    vals = np.array([1.0, 1.00390625, 1.001953125, 1.005859375, -3.0, 0.0], dtype=np.float32)
    # 1 + 2^-8 is exactly halfway between 1 and 1 + 2^-7: ties to even -> 1
    # 1 + 3*2^-9 rounds up to 1 + 2^-7; 1 + 2^-9 rounds down to 1 (ulp at 1 is 2^-7)
    exp = np.array([1.0, 1.0, 1.0, 1.0078125, -3.0, 0.0])
    got = M.bf16_round(vals)
    assert np.array_equal(got[[0, 4, 5]], exp[[0, 4, 5]])
    assert got[1] == 1.0 and got[2] == 1.0 and got[3] == 1.0078125


def test_gelu_grad_matches_differences():
    x = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(x))) < 1e-8


@pytest.mark.parametrize("dims", [[5, 8, 3], [7, 16, 16, 4]])
def test_gradient_matches_finite_differences(dims):
    rng = np.random.default_rng(1)
    flat = M.init_params(dims, seed=3).astype(np.float64)
    flat += 0.05 * rng.standard_normal(flat.shape)  # nonzero biases too
    X = rng.standard_normal((9, dims[0]))
    T = rng.standard_normal((9, dims[-1]))
    loss, g = M.loss_and_grad(flat, dims, X, T)
    h = 1e-6
    fd = np.empty_like(flat)
    for i in range(flat.size):
        e = np.zeros_like(flat)
        e[i] = h
        fd[i] = (M.loss_and_grad(flat + e, dims, X, T)[0] - M.loss_and_grad(flat - e, dims, X, T)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_adamw_first_step_closed_form():
    dims = [3, 2]
    flat = np.array([0.5, -0.25, 1.0, 2.0, -1.0, 0.75, 0.1, -0.2])
    g = np.array([0.3, -0.1, 0.2, 2.0, -5.0, 1e-3, 0.4, -0.4])
    lr, wd = 1e-2, 0.1
    new, m, v = M.adamw_step(flat, g, np.zeros_like(g), np.zeros_like(g), 1, dims, lr, wd=wd, eps=0.0)
    decay = np.array([wd] * 6 + [0.0, 0.0])
    expect = flat * (1 - lr * decay) - lr * np.sign(g)
    assert np.allclose(new, expect, rtol=0, atol=1e-15)
