"""The FP64 restatement of the seeder network (oracle/mlp.py) against
central finite differences, and its BF16 rounding against values the format
fixes."""
from __future__ import annotations

import numpy as np

from oracle import mlp


def test_bf16_rounding():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -2.5, 3.14159265, 1e-40, 65504.0], dtype=np.float64)
    want = np.array([1.0, 1.0, 1 + 2 ** -6, -2.5, 3.140625, 0.0, 65536.0])
    got = mlp.bf16(x)
    np.testing.assert_array_equal(got[:5], want[:5])
    assert abs(got[5]) < 1e-38 and got[6] == want[6]


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    F, H, NH, O, B = 6, 5, 2, 3, 7
    dims = mlp.layer_dims(F, O, H, NH)
    n = sum(a * b + b for a, b in dims)
    p = rng.normal(0, 0.5, n)
    x, y = rng.normal(size=(B, F)), rng.normal(size=(B, O))
    loss, g = mlp.loss_grad(p, x, y, O, dims)
    h = 1e-6
    # only parameters that influence the real outputs have nonzero gradients
    for k in rng.choice(n, 60, replace=False):
        e = np.zeros(n); e[k] = h
        fd = (mlp.loss_grad(p + e, x, y, O, dims)[0] - mlp.loss_grad(p - e, x, y, O, dims)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), k
    # the padded output columns get no gradient through the loss
    W_last_off = sum(a * b + b for a, b in dims[:-1])
    Wl = g[W_last_off:W_last_off + dims[-1][0] * dims[-1][1]].reshape(dims[-1])
    assert np.all(Wl[:, O:] == 0)


def test_gelu_derivative():
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    np.testing.assert_allclose((mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h), mlp.gelu_grad(z), atol=1e-8)
