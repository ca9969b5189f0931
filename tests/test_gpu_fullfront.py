"""The whole-front-in-LDS factor kernel (k_factor_full) against the panel
kernel (k_factor_large): every bucket whose fronts fit is switched to a
whole-front launch class by hand (nr_batch_set_launch_class), and the solves
must give the same voltages, iteration counts and KCL bit for bit, in FP64
and on the FP32 path, since every front entry receives the same operations in
the same order in both kernels."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu

FULL_CLASSES = [(4, 256), (8, 256), (16, 256), (32, 256), (64, 64)]


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def force_full(bt, prec):
    from karam_amd import nr
    chosen = []
    for b in range(8):
        for team, block in FULL_CLASSES:
            try:
                bt.set_launch_class(b, team, block, full=1, precision=prec)
                chosen.append((b, team))
                break
            except nr.NrError:
                continue
    return chosen


def circuits():
    return {
        "rdmesh32": (netlists.rd_mesh(32, seed=4), 64, {}),
        "rdmesh64": (netlists.rd_mesh(64, seed=1), 48, {}),
        "hightw": (netlists.random_logic(2000, seed=3, spread=32, window=6, width=64), 4,
                   dict(sigma_vto=0.03, sigma_kp=0.03)),
    }


@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
@pytest.mark.parametrize("name", ["rdmesh32", "rdmesh64", "hightw"])
def test_full_front_equals_panel(ctx, name, prec):
    from karam_amd import nr
    nl, B, kw = circuits()[name]
    f = netlists.mc_fields(nl, B, seed=31, **kw)
    ref = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    ref.set_fields(f)
    ref.solve_dc(precision=prec)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    bt.set_fields(f)
    chosen = force_full(bt, prec)
    assert chosen, "no bucket took a whole-front class"
    bt.solve_dc(precision=prec, record_timing=2)
    assert any(k.startswith("k_factor_full") for k in bt.kernel_stats())
    np.testing.assert_array_equal(bt.voltages(), ref.voltages())
    s0, s1 = ref.stats(), bt.stats()
    for k in ("iterations", "status", "kcl"):
        np.testing.assert_array_equal(s1[k], s0[k], err_msg=k)


def test_unknown_and_oversized_classes(ctx):
    from karam_amd import nr
    nl = netlists.rd_mesh(16, seed=2)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), 4)
    with pytest.raises(nr.NrError):
        bt.set_launch_class(0, 3, 256, full=1)            # no 3-lane class
    with pytest.raises(nr.NrError):
        bt.set_launch_class(9, 4, 256)                    # no bucket 9
    # bucket 6 (orders 65-96) in 4-lane whole-front teams: 64 teams x 96 x 97
    # doubles exceed the LDS a block holds (only refused when the bucket has fronts)
    T = nr.Topology(netlists.rd_mesh(64, seed=1), ctx)
    big = nr.Batch(ctx, T, 4)
    with pytest.raises(nr.NrError):
        big.set_launch_class(6, 4, 256, full=1)
