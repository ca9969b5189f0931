"""The dense pivoting LU's blocked elimination order (k_dense) against the
oracle's restatement of solveLU_Pivoted (math/linalg.h:742-854), on the CPU.

k_dense factors a panel of w columns step by step (pivot search, the swap
of the panel part of the two rows and of the rhs, multipliers, the panel's
own updates), then applies the panel's row swaps to the rest of each row in
step order, forms the pivot rows' entries right of the panel from the
earlier pivot rows in step order, and gives every trailing entry the
panel's w updates in step order.  This restates that order in numpy (one
IEEE operation per element, no fused multiply-add) and checks that it
reproduces the unblocked routine bit for bit for panel widths 1-16 and
varied panel capacities, on random systems with ties and singular cases.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def blocked_lu_solve(A, b, cap, nb, reverse_trailing=False):
    n = len(b)
    M = A.copy()
    sy = b.copy()
    i0 = 0
    while i0 < n:
        rows = n - i0
        w = min(nb, rows, cap // rows)
        pan = M[i0:, i0:i0 + w].copy()
        piv = []
        for k in range(w):
            bv, bk = 0.0, None
            for r in range(k, rows):
                v = abs(pan[r, k])
                if v > bv:
                    bv, bk = v, r
            if bv < 1e-25:
                return None
            p = bk
            piv.append(i0 + p)
            if p != k:
                pan[[k, p]] = pan[[p, k]]
                sy[i0 + k], sy[i0 + p] = sy[i0 + p], sy[i0 + k]
            pan[k + 1:, k] = pan[k + 1:, k] / pan[k, k]
            for r in range(k + 1, rows):
                l = pan[r, k]
                if l != 0.0:
                    pan[r, k + 1:w] = pan[r, k + 1:w] - l * pan[k, k + 1:w]
        M[i0:, i0:i0 + w] = pan
        other = np.r_[0:i0, i0 + w:n]
        for k in range(w):
            q = piv[k]
            if q != i0 + k:
                t = M[i0 + k, other].copy()
                M[i0 + k, other] = M[q, other]
                M[q, other] = t
        c0 = i0 + w
        for k in range(1, w):
            for q in range(k):
                l = pan[k, q]
                if l != 0.0:
                    M[i0 + k, c0:] = M[i0 + k, c0:] - l * M[i0 + q, c0:]
        for r in range(w, rows):
            for k in (reversed(range(w)) if reverse_trailing else range(w)):
                l = pan[r, k]
                if l != 0.0:
                    M[i0 + r, c0:] = M[i0 + r, c0:] - l * M[i0 + k, c0:]
        i0 += w
    y = sy.copy()
    for i in range(n):
        acc = y[i]
        for j in range(i):
            acc = acc - M[i, j] * y[j]
        y[i] = acc
    x = np.zeros(n)
    for i in range(n - 1, -1, -1):
        xi = y[i]
        for j in range(i + 1, n):
            xi = xi - M[i, j] * x[j]
        x[i] = xi / M[i, i]
    return x


def _csr(A):
    n = A.shape[0]
    rp = np.zeros(n + 1, np.int32)
    ci, vals = [], []
    for i in range(n):
        nz = np.flatnonzero(A[i])
        ci += list(nz)
        vals += list(A[i, nz])
        rp[i + 1] = len(ci)
    return rp, np.array(ci, np.int32), np.array(vals)


@pytest.mark.parametrize("trial", range(16))
def test_blocked_order_is_the_reference_lu(trial):
    rng = np.random.default_rng(100 + trial)
    n = int(rng.integers(3, 60))
    dens = rng.uniform(0.05, 0.6)
    A = np.where(rng.random((n, n)) < dens, rng.normal(size=(n, n)) * 10.0 ** rng.integers(-6, 6, size=(n, n)), 0.0)
    A[rng.random(n) < 0.5, :] *= 1e3
    if trial % 4 == 0:              # ties in the pivot search: small integers
        A = np.where(A != 0, np.round(A / np.abs(A).max() * 4), 0.0)
    A += np.diag(np.where(rng.random(n) < 0.7, rng.normal(size=n), 0.0))
    b = rng.normal(size=n)
    rc, xo = O.lu_dense(n, *_csr(A), b)
    cap = int(rng.integers(n, 4 * n))
    nb = int(rng.choice([1, 2, 3, 16]))
    xb = blocked_lu_solve(A, b, cap, nb)
    if rc != 0:
        assert xb is None
    else:
        assert xb is not None
        assert np.array_equal(xb.view(np.int64), xo.view(np.int64))


def test_a_wrong_order_is_caught():
    """The same restatement with each trailing entry's panel updates applied
    in reverse step order misses the oracle's bits on a generic system, so
    the comparison above is sensitive to the order it claims to keep."""
    rng = np.random.default_rng(7)
    n = 40
    A = rng.normal(size=(n, n)) + np.diag(rng.normal(size=n))
    b = rng.normal(size=n)
    rc, xo = O.lu_dense(n, *_csr(A), b)
    assert rc == 0
    assert np.array_equal(blocked_lu_solve(A, b, 8 * n, 8).view(np.int64), xo.view(np.int64))
    xr = blocked_lu_solve(A, b, 8 * n, 8, reverse_trailing=True)
    assert not np.array_equal(xr.view(np.int64), xo.view(np.int64))
