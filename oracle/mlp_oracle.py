"""CPU restatement of the surrogate network -- TEST INFRASTRUCTURE ONLY.

Float64 forward pass, backward pass and AdamW step of the MLP that
cosenu_amd/csrc/mlp.hip trains on the GPU: y = x W + b per layer, exact
(erf) GELU between layers, mean-squared-error loss over all outputs, AdamW
with decoupled weight decay (none on biases).  With ``bf16=True`` the values
the GPU rounds to bfloat16 before a matrix product (inputs, weights,
activations, back-propagated gradients) are rounded the same way here, so the
remaining GPU/CPU difference is only FP32-vs-FP64 accumulation.

There is no reference counterpart (the reference has no learned part); this
restatement is pinned by central finite differences of its own loss
(tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

try:
    from scipy.special import erf as _erf
except Exception:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def bf16_round(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def unflatten(flat, dims):
    out, o = [], 0
    for l in range(len(dims) - 1):
        k, n = dims[l], dims[l + 1]
        W = np.asarray(flat[o:o + k * n], dtype=np.float64).reshape(k, n)
        o += k * n
        b = np.asarray(flat[o:o + n], dtype=np.float64)
        o += n
        out.append((W, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def init_params(dims, seed=0):
    """He-uniform weights, zero biases (float32 values)."""
    rng = np.random.default_rng(seed)
    layers = []
    for l in range(len(dims) - 1):
        k, n = dims[l], dims[l + 1]
        lim = math.sqrt(6.0 / k)
        layers.append((rng.uniform(-lim, lim, (k, n)).astype(np.float32).astype(np.float64),
                       np.zeros(n)))
    return flatten(layers).astype(np.float32)


def forward(flat, dims, X, bf16=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(flat, dims)
    a = r(X)
    cache = [a]
    Zs = []
    for l, (W, b) in enumerate(layers):
        z = a @ r(W) + b
        Zs.append(z)
        if l < len(layers) - 1:
            a = r(gelu(z))
            cache.append(a)
    return Zs[-1], (cache, Zs)


def loss_and_grad(flat, dims, X, T, bf16=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(flat, dims)
    Y, (acts, Zs) = forward(flat, dims, X, bf16)
    n = Y.size
    d = Y - T
    loss = float(np.sum(d * d) / n)
    dZ = 2.0 * d / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        dZr = r(dZ)
        gW = acts[l].T @ dZr
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        dA = dZr @ r(W).T
        dZ = dA * gelu_grad(Zs[l - 1])
    return loss, flatten(grads)


def adamw_step(flat, grad, m, v, t, dims, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW step (bias-corrected moments, decoupled decay, none on
    biases).  Returns new (flat, m, v)."""
    flat = np.asarray(flat, dtype=np.float64).copy()
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    bc1, bc2 = 1 - b1 ** t, 1 - b2 ** t
    decay = np.zeros_like(flat)
    o = 0
    for l in range(len(dims) - 1):
        k, n = dims[l], dims[l + 1]
        decay[o:o + k * n] = wd
        o += k * n + n
    flat = flat * (1 - lr * decay)
    flat -= lr * (m / bc1) / (np.sqrt(v / bc2) + eps)
    return flat, m, v
