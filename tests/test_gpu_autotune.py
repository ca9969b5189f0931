"""The autotuner (nr_batch_autotune): it returns a launch shape for every
front-order bucket that holds large fronts, and solves after tuning are
bit-identical to solves before it (the elimination's operations and their
order do not depend on the team size)."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
def test_autotune_keeps_results(ctx, prec):
    from karam_amd import nr
    nl = netlists.rd_mesh(40, seed=3)
    B = 192
    f = netlists.mc_fields(nl, B, seed=23)
    T = nr.Topology(nl, ctx)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    bt.solve_dc(precision=prec)
    V, st = bt.voltages(), bt.stats()
    tune = bt.autotune(precision=prec, reps=2)
    assert tune["ms_before"] > 0 and tune["ms_after"] > 0
    used = [b for b in tune["buckets"] if b["fronts"] > 0]
    assert used and sum(b["fronts"] for b in tune["buckets"]) > 0
    for b in used:
        assert b["team"] in (4, 8, 16, 32, 64, 128, 256) and b["block"] % b["team"] == 0
    res = bt.solve_dc(precision=prec)
    assert res["n_converged"] == B
    np.testing.assert_array_equal(bt.voltages(), V)
    np.testing.assert_array_equal(bt.stats()["iterations"], st["iterations"])
    # a second batch of the same topology uses the tuned shapes too
    b2 = nr.Batch(ctx, T, 8)
    b2.set_fields({k: v[:8] for k, v in f.items()})
    b2.solve_dc(precision=prec)
    np.testing.assert_array_equal(b2.voltages(), V[:8])


def test_autotune_rejects_bad_precision(ctx):
    from karam_amd import nr
    nl = netlists.rd_mesh(8, seed=1)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), 4)
    with pytest.raises(nr.NrError):
        bt.autotune(precision=7)
