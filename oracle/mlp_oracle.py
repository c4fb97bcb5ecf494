"""CPU FP64 oracle of the learned initialiser's MLP (4*nl -> 256 -> 256 ->
256 -> 10, tanh-form GELU), forward and backward.

TEST INFRASTRUCTURE ONLY (imported by tests/): checks the device BF16-MFMA
training kernels.  With ``bf16=True`` the operands are rounded to bfloat16 at
the points where the device rounds them (inputs, weights, hidden
activations, the output gradient and the hidden-layer gradients), so the
remaining difference is FP32 vs FP64 accumulation.  Pinned by central
finite differences of the loss (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

H = 256
OUT = 10


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


K_GELU = 2.0 * np.sqrt(2.0 / np.pi)


def gelu(z):
    """GELU in its tanh form, z * sigmoid(k (z + 0.044715 z^3)), k = 2 sqrt(2/pi):
    the activation the device network uses (within 5e-4 of z Phi(z))."""
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(over="ignore"):
        return z / (1.0 + np.exp(-K_GELU * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(over="ignore"):
        s = 1.0 / (1.0 + np.exp(-K_GELU * (z + 0.044715 * z ** 3)))
    return s + z * s * (1.0 - s) * K_GELU * (1.0 + 3 * 0.044715 * z * z)


def gelu_erf(z):
    """z Phi(z), the exact GELU the tanh form approximates."""
    return 0.5 * z * (1.0 + erf(np.asarray(z, dtype=np.float64) / np.sqrt(2.0)))


def unpack(params, nin):
    """Exported flat order: per layer W [din][dout] (row-major) then b [dout]."""
    dims = [nin, H, H, H, OUT]
    out, o = [], 0
    for l in range(4):
        din, dout = dims[l], dims[l + 1]
        W = np.asarray(params[o:o + din * dout], dtype=np.float64).reshape(din, dout)
        o += din * dout
        b = np.asarray(params[o:o + dout], dtype=np.float64)
        o += dout
        out.append((W, b))
    assert o == len(params)
    return out


def num_params(nin):
    return nin * H + H + 2 * (H * H + H) + H * OUT + OUT


def loss_and_grad(params, x, y, bf=True):
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, x.shape[1])
    B = x.shape[0]
    acts = [r(x)]
    zs = []
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ r(W) + b
        zs.append(z)
        if l < 3:
            acts.append(r(gelu(z)))
    diff = zs[-1] - y
    loss = float(np.mean(diff * diff))
    dz = r(2.0 * diff / (B * OUT))
    grads = [None] * 4
    for l in range(3, -1, -1):
        W, b = layers[l]
        grads[l] = (acts[l].T @ dz, dz.sum(axis=0))
        if l > 0:
            dz = r((dz @ r(W).T) * gelu_grad(zs[l - 1]))
    flat = np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])
    return loss, flat


def init_params(nin, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    dims = [nin, H, H, H, OUT]
    parts = []
    for l in range(4):
        a = np.sqrt(6.0 / dims[l]) * scale
        parts.append(rng.uniform(-a, a, size=dims[l] * dims[l + 1]))
        parts.append(rng.uniform(-0.1, 0.1, size=dims[l + 1]))
    return np.concatenate(parts)
