"""CPU restatement of the learned initializer's network, for the tests only.

A multilayer perceptron with nHiddenLayers hidden layers of nHidden units, the
exact GELU activation GELU(z) = z Phi(z) = z (1 + erf(z / sqrt 2)) / 2, a linear
output layer, and the mean squared error L = sum (y - t)^2 / (rows * nOut).
Parameters are ordered W_0, b_0, W_1, b_1, ..., W_L, b_L with W_l stored
row-major as [in][out] (y = x W + b), the layout of rpwa_mlp_get_params.
Everything is float64; forward and backward are written out by hand (the
chain rule through each layer), so the tests can compare the GPU kernels'
FP32 gradients with these and pin these against finite differences.

This restates no reference code: the network has no counterpart in ROOTPWA.
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover - scipy is part of the image
    _erf = np.vectorize(math.erf)

SQRT2 = math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / SQRT2))


def gelu_grad(z):
    """d GELU / dz = Phi(z) + z phi(z)."""
    return 0.5 * (1.0 + _erf(z / SQRT2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def shapes(nIn, nHidden, nHiddenLayers, nOut):
    dims = [nIn] + [nHidden] * nHiddenLayers + [nOut]
    return [(dims[l], dims[l + 1]) for l in range(len(dims) - 1)]


def unpack(params, nIn, nHidden, nHiddenLayers, nOut):
    out, off = [], 0
    for (i, o) in shapes(nIn, nHidden, nHiddenLayers, nOut):
        W = np.asarray(params[off:off + i * o], dtype=np.float64).reshape(i, o)
        off += i * o
        b = np.asarray(params[off:off + o], dtype=np.float64)
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def forward(layers, x):
    """Returns the output and the per-layer (input, pre-activation) pairs."""
    h = np.asarray(x, dtype=np.float64)
    cache = []
    for l, (W, b) in enumerate(layers):
        z = h @ W + b
        cache.append((h, z))
        h = gelu(z) if l + 1 < len(layers) else z
    return h, cache


def loss_grad(layers, x, t):
    """MSE loss and its gradient with respect to every parameter (flat, in parameter order)."""
    y, cache = forward(layers, x)
    t = np.asarray(t, dtype=np.float64)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = 2.0 * (y - t) / n
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        h, z = cache[l]
        W, _ = layers[l]
        grads[l] = (h.T @ dz, dz.sum(axis=0))
        if l > 0:
            dh = dz @ W.T
            dz = dh * gelu_grad(cache[l - 1][1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def loss_only(layers, x, t):
    y, _ = forward(layers, x)
    return float(np.sum((y - np.asarray(t, dtype=np.float64)) ** 2) / y.size)
