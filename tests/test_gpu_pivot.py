"""Long CMOS inverter chains, where the multifrontal LU without pivoting grows
unstable (each stage's gm exceeds its output conductance by the stage gain,
so eliminating the chain in Min-Fill order multiplies the supply column by
that gain at every stage).

- The default route (NR_LU_AUTO: multifrontal, then the relative residual of
  every solve, the dense partial-pivoting LU for the solves above the bound)
  against the oracle's same route: bit-identical, level-1 MOSFETs involve no
  exp / log.
- The same against the reference's own LU, the dense partial-pivoting
  solveLU_Pivoted (math/linalg.h:742-854) that solveDC runs (solvers/
  circuitsim_solver.cpp:291-295), restated by the oracle (orc_lu_dense) and
  stored in tests/golden/chain_dense.npz: every instance converges where the
  reference's does, at its voltages, and fails where it fails (at 1000
  stages with VTO sigma 5 mV, KP sigma 1 %, two of four instances do not
  converge in the reference either).  Iteration counts are printed, not held
  to +-1: on these chains the reference's own count moves by up to 10 when
  one MOSFET width in two changes by 1e-16 relative (its Newton iteration
  at the third GMIN step is that sensitive to rounding), so only the same
  arithmetic reproduces it, which is what NR_LU_DENSE is for.
- NR_LU_DENSE against the oracle's dense LU: bit-identical, iterations equal.
"""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

from conftest import load_golden

pytestmark = pytest.mark.gpu

B = 4
SEED = 7      # as tests/golden/make_chain_dense.py
SIGMA = {150: (0.03, 0.03), 300: (0.03, 0.03), 500: (0.03, 0.03), 1000: (0.005, 0.01)}


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


@pytest.fixture(scope="module")
def golden():
    return load_golden("chain_dense")


def _batch(ctx, L):
    from karam_amd import nr
    nl = netlists.inverter_chain(L)
    f = netlists.mc_fields(nl, B, seed=SEED, sigma_vto=SIGMA[L][0], sigma_kp=SIGMA[L][1])
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    bt.set_fields(f)
    return nl, f, bt


def _err_over_tol(v, ref):
    return float((np.abs(v - ref) / (1e-9 * np.abs(ref) + 1e-12)).max())


@pytest.mark.parametrize("L", [150, 300, 500, 1000])
def test_auto_route_converges_where_the_reference_does(ctx, golden, L):
    from oracle import oracle as O
    nl, f, bt = _batch(ctx, L)
    res = bt.solve_dc()
    V, st = bt.voltages(), bt.stats()
    print(f"inv{L}: GPU AUTO iterations {st['iterations'].tolist()} (dense-LU fallbacks "
          f"{st['lu_fallbacks'].tolist()}), reference LU {golden[f'it{L}'].tolist()}", flush=True)
    ok = golden[f"st{L}"] == _abi.NR_STATUS_CONVERGED
    assert ok.sum() >= 2
    np.testing.assert_array_equal(st["status"], golden[f"st{L}"])
    assert res["n_converged"] == ok.sum() and (st["kcl"][ok] <= 1e-9).all()
    assert (st["lu_fallbacks"] >= 1).all()
    assert _err_over_tol(V[ok], golden[f"v{L}"][ok]) <= 1.0
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), f, B, lu=O.LU_AUTO, threads=B)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(st["iterations"], so["iterations"])
    np.testing.assert_array_equal(st["lu_fallbacks"], so["lu_fallbacks"])


@pytest.mark.parametrize("L", [150, 300])
def test_auto_route_fp32_path(ctx, golden, L):
    """The FP32 path's float solve is redone in FP64 by the dense LU when its
    relative residual exceeds 1e-3; bit-identical to the oracle's FP32
    restatement, within tolerance of the reference's FP64 answer."""
    from oracle import oracle as O
    nl, f, bt = _batch(ctx, L)
    res = bt.solve_dc(precision=_abi.NR_PREC_FP32_REFINE)
    V, st = bt.voltages(), bt.stats()
    assert res["n_converged"] == B and (st["kcl"] <= 1e-9).all()
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), f, B, lu=O.LU_AUTO, threads=B, precision=O.PREC_FP32_REFINE)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(st["iterations"], so["iterations"])
    np.testing.assert_array_equal(st["lu_fallbacks"], so["lu_fallbacks"])
    assert _err_over_tol(V, golden[f"v{L}"]) <= 1.0


@pytest.mark.parametrize("L", [150, 300])
def test_dense_route_is_the_reference_lu(ctx, golden, L):
    nl, f, bt = _batch(ctx, L)
    res = bt.solve_dc(lu=_abi.NR_LU_DENSE)
    V, st = bt.voltages(), bt.stats()
    assert res["n_converged"] == B
    np.testing.assert_array_equal(V, golden[f"v{L}"])
    np.testing.assert_array_equal(st["iterations"], golden[f"it{L}"])
    np.testing.assert_array_equal(st["ref_iterations"], golden[f"ref{L}"])
    assert (st["lu_fallbacks"] == 0).all()


def test_multifrontal_route_alone_fails_like_the_oracle(ctx):
    """Without the fallback the chain fails, exactly as the oracle's
    multifrontal restatement does."""
    from oracle import oracle as O
    nl, f, bt = _batch(ctx, 300)
    bt.solve_dc(lu=_abi.NR_LU_MULTIFRONTAL)
    st = bt.stats()
    _, so = O.solve_dc_batch(nl, O.Topology(nl), f, B, lu=O.LU_MULTIFRONTAL, threads=B)
    assert (so["status"] != _abi.NR_STATUS_CONVERGED).all()
    np.testing.assert_array_equal(st["status"], so["status"])
    np.testing.assert_array_equal(st["iterations"], so["iterations"])


def test_dense_route_refuses_large_circuits(ctx):
    from karam_amd import nr
    nl = netlists.rd_mesh(65, seed=3)            # 4225 nodes > NR_DENSE_MAX_N
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), 2)
    with pytest.raises(nr.NrError, match="NR_LU_DENSE"):
        bt.solve_dc(lu=_abi.NR_LU_DENSE)
    r = bt.solve_dc()                            # AUTO: the multifrontal solve alone
    assert r["n_converged"] == 2
    assert (bt.stats()["lu_fallbacks"] == 0).all()
