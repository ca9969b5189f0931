"""High-treewidth circuits: random NAND / INV logic whose gates draw their
inputs from 65 slots of the previous six levels (no locality), 2000 nodes,
treewidth 214.  The reference routes such circuits away from its sparse LU
(treewidth above 50; solvers/circuitsim_solver.cpp:288-351): to its tensor
contraction (<= 80), its Schur domain decomposition (<= 100) or PCG with a
dense-LU fallback.  Here the same multifrontal factorisation runs at any
treewidth; its largest fronts (order 215) are factored with their panels in
HBM (in FP64).  Results against the oracle: bit-identical in FP64 (exp-free circuit),
to the refinement tolerance in FP32."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


def circuit():
    return netlists.random_logic(2000, seed=3, spread=32, window=6, width=64)


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


@pytest.fixture(scope="module")
def oracle_result():
    from oracle import oracle as O
    nl = circuit()
    B = 4
    f = netlists.mc_fields(nl, B, seed=5, sigma_vto=0.03, sigma_kp=0.03)
    V, st = O.solve_dc_batch(nl, O.Topology(nl), f, B, threads=4)
    return nl, f, B, V, st


@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
def test_high_treewidth_dc(ctx, oracle_result, prec):
    from karam_amd import nr
    nl, f, B, Vo, so = oracle_result
    T = nr.Topology(nl, ctx)
    assert T.info["treewidth"] > 100 and T.info["max_front"] > 200
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    res = bt.solve_dc(precision=prec, record_timing=2)
    if prec == _abi.NR_PREC_FP64:      # the largest FP64 panels exceed the LDS a block holds
        assert any(k.startswith("k_factor_large<256, 256, true") for k in bt.kernel_stats())
    V, st = bt.voltages(), bt.stats()
    assert np.all(so["status"] == 0) and res["n_converged"] == B
    if prec == _abi.NR_PREC_FP64:
        np.testing.assert_array_equal(V, Vo)
        np.testing.assert_array_equal(st["iterations"], so["iterations"])
        np.testing.assert_array_equal(st["kcl"], so["kcl"])
    else:
        assert st["kcl"].max() <= 1e-9
        assert np.all(np.abs(V - Vo) <= 1e-6 * np.abs(Vo) + 1e-9)
