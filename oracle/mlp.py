"""FP64 restatement of the seeder network for checking its GPU kernels.

TEST INFRASTRUCTURE ONLY — the checker of karam_amd's learned seeder
(MLP: h_l = GELU(h_{l-1} W_l + b_l) for the hidden layers, linear output,
mean squared error).  The flat parameter layout is the library's: per layer
W [in x out] row-major, then b [out].  With bf16=True the operands of every
matrix product are rounded to BF16 as the GPU rounds them (inputs, weights,
hidden activations and back-propagated gradients), everything else FP64.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def layer_dims(n_features, n_out, hidden, n_hidden):
    n_out_p = (n_out + 63) // 64 * 64
    dims = [n_features] + [hidden] * n_hidden + [n_out_p]
    return list(zip(dims[:-1], dims[1:]))


def unflatten(p, dims):
    out, o = [], 0
    for din, dout in dims:
        W = p[o:o + din * dout].reshape(din, dout); o += din * dout
        b = p[o:o + dout]; o += dout
        out.append((W, b))
    return out


def bf16(x):
    """Round to BF16 (nearest even) and back, in float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def loss_grad(p, x, y, n_out, dims, round_bf16=False):
    """Mean squared error over the n_out real outputs and its gradient."""
    q = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unflatten(np.asarray(p, np.float64), dims)
    B = x.shape[0]
    n_out_p = dims[-1][1]
    yp = np.zeros((B, n_out_p)); yp[:, :n_out] = y
    hs, zs = [q(x)], []
    for li, (W, b) in enumerate(layers):
        z = hs[-1] @ q(W) + b
        zs.append(z)
        if li < len(layers) - 1:
            hs.append(q(gelu(z)))
    yhat = zs[-1]
    d = (yhat - yp)[:, :n_out]
    loss = (d ** 2).sum() / (B * n_out)
    dz = np.zeros_like(yhat); dz[:, :n_out] = 2.0 * d / (B * n_out)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        dzq = q(dz)
        grads.append((hs[li].T @ dzq, dz.sum(0)))
        if li > 0:
            dh = dzq @ q(W).T
            dz = dh * gelu_grad(zs[li - 1])
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
