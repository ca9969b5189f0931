"""nr_eval_mosfets_noalloc (staging in the library's fixed per-GPU scratch, no
device allocation) against nr_eval_mosfets (staging allocated per call),
both through the C ABI: two calls in a row, the second smaller, each
bit-identical to the allocating entry point; a call beyond the scratch is
refused with NR_ERR_INVALID."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from karam_amd import _abi

pytestmark = pytest.mark.gpu

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


def _inputs(m, nv, seed):
    rng = np.random.default_rng(seed)
    d = rng.integers(0, nv + 1, m).astype(np.int32)
    g = rng.integers(0, nv + 1, m).astype(np.int32)
    s = rng.integers(0, nv + 1, m).astype(np.int32)
    pm = rng.integers(0, 2, m).astype(np.int32)
    W = rng.uniform(1e-6, 8e-6, m)
    L = rng.uniform(0.5e-6, 2e-6, m)
    Kp = rng.uniform(40e-6, 120e-6, m)
    Vth = np.where(pm == 1, -1.0, 1.0) * rng.uniform(0.3, 0.9, m)
    lam = rng.uniform(0.0, 0.08, m)
    v = rng.uniform(-0.5, 2.3, nv)
    return d, g, s, W, L, Kp, Vth, lam, pm, v


def _call(fn, ctx, args):
    d, g, s, W, L, Kp, Vth, lam, pm, v = args
    m = len(d)
    outs = [np.full(m, np.nan) for _ in range(5)]
    p32 = lambda a: a.ctypes.data_as(_i32p)   # noqa: E731
    p64 = lambda a: a.ctypes.data_as(_f64p)   # noqa: E731
    rc = fn(ctx.h, m, p32(d), p32(g), p32(s), p64(W), p64(L), p64(Kp), p64(Vth), p64(lam), p32(pm), p64(v), len(v),
            *[p64(o) for o in outs])
    return rc, outs


def test_noalloc_equals_the_allocating_slot_on_two_calls():
    from karam_amd import nr
    ctx = nr.Context(0)
    L = nr.lib()
    for m, nv, seed in ((60000, 20000, 1), (1000, 300, 2)):      # the second call smaller
        args = _inputs(m, nv, seed)
        rc_a, ref = _call(L.nr_eval_mosfets, ctx, args)
        rc_b, got = _call(L.nr_eval_mosfets_noalloc, ctx, args)
        assert rc_a == rc_b == _abi.NR_OK
        for a, b in zip(ref, got):
            assert not np.isnan(b).any()
            np.testing.assert_array_equal(a, b)


def test_noalloc_refuses_calls_beyond_its_scratch():
    from karam_amd import nr
    ctx = nr.Context(0)
    args = _inputs(90000, 90000, 3)
    rc, _ = _call(nr.lib().nr_eval_mosfets_noalloc, ctx, args)
    assert rc == _abi.NR_ERR_INVALID
    assert b"scratch" in nr.lib().nr_last_error(ctx.h)
    # still usable afterwards
    rc, outs = _call(nr.lib().nr_eval_mosfets_noalloc, ctx, _inputs(10, 5, 4))
    assert rc == _abi.NR_OK and not np.isnan(outs[0]).any()
