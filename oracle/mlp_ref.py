"""FP64 CPU reference of the surrogate MLP (TEST INFRASTRUCTURE ONLY).

Forward pass and mean-squared-error gradient of the network the engine
trains on the GPU (lcurve_re_amd/csrc/k_mlp.hip): SiLU hidden layers, a
linear output layer, weights row-major [fan_in x fan_out], parameters laid
out layer by layer as W_l then b_l.

With bf16=True the values are rounded to bfloat16 (round to nearest even)
at the points where the GPU kernels round them -- the network input, the
weights, the hidden activations and the back-propagated gradients that feed
a matrix product -- and everything else is FP64, so the GPU result differs
from this one only by FP32 accumulation.  The gradient itself is pinned by
central finite differences in tests/test_surrogate.py.
"""
import numpy as np


def bf16(a):
    """Rounds finite float values to bfloat16 (nearest even), as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, widths):
    out, o = [], 0
    for l in range(1, len(widths)):
        k, n = widths[l - 1], widths[l]
        W = np.asarray(params[o:o + k * n], np.float64).reshape(k, n)
        o += k * n
        b = np.asarray(params[o:o + n], np.float64)
        o += n
        out.append((W, b))
    assert o == len(params)
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, widths, x, bf=False):
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, widths)
    a = r(x)
    zs, acts = [], [a]
    for l, (W, b) in enumerate(layers):
        z = a @ r(W) + b
        zs.append(z)
        if l < len(layers) - 1:
            a = r(_silu(z))
            acts.append(a)
        else:
            a = z
    return a, zs, acts


def grad(params, widths, x, t, bf=False):
    """(mean squared error, gradient) over the rows of x."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, widths)
    y, zs, acts = forward(params, widths, x, bf)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    g = 2.0 * (y - t) / n
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        gq = r(g)
        dW = acts[l].T @ gq
        db = g.sum(axis=0)
        grads.append((dW, db))
        if l > 0:
            da = gq @ r(W).T
            g = da * _silu_grad(zs[l - 1])
    grads.reverse()
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat
