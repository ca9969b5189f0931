"""The whole-phase large-front kernels (k_factor_persist, k_backsolve_persist,
nr_batch_set_persistent) and the level fan-out (nr_batch_set_fanout)
against the per-level launches on one stream: one block per
instance walks every elimination-tree level of the large fronts, so the
solves must give the same voltages, iteration counts, statuses and KCL bit
for bit, in FP64 and on the FP32 path, eagerly and replayed from hipGraphs,
for DC operating points and transients (every front receives the same
operations in the same order whatever team runs it).  Fronts whose panel
exceeds a team's LDS share take their panel in HBM (the 64 x 64 mesh has
such fronts)."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def circuits():
    return {
        "rdmesh64": (netlists.rd_mesh(64, seed=1), 64, {}),
        "logic2000": (netlists.random_logic(2000, seed=1, width=64, p_inv=0.3, p_nand3=0.1), 96,
                      dict(sigma_vto=0.03, sigma_kp=0.03)),
        "hightw": (netlists.random_logic(2000, seed=3, spread=32, window=6, width=64), 8,
                   dict(sigma_vto=0.03, sigma_kp=0.03)),
    }


def _pair(ctx, nl, B, f, transient=False):
    from karam_amd import nr
    ref = nr.Batch(ctx, nr.Topology(nl, ctx, transient=transient), B)
    ref.set_fields(f)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx, transient=transient), B)
    bt.set_fields(f)
    return ref, bt


def _same(a, b):
    np.testing.assert_array_equal(b.voltages(), a.voltages())
    s0, s1 = a.stats(), b.stats()
    for k in ("iterations", "status", "kcl"):
        np.testing.assert_array_equal(s1[k], s0[k], err_msg=k)


@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
@pytest.mark.parametrize("name", ["rdmesh64", "logic2000", "hightw"])
def test_persistent_equals_level_launches(ctx, name, prec):
    nl, B, kw = circuits()[name]
    f = netlists.mc_fields(nl, B, seed=41, **kw)
    ref, bt = _pair(ctx, nl, B, f)
    print(f"persistent {name} prec={prec} B={B}", flush=True)
    ref.solve_dc(precision=prec)
    bt.set_persistent(True, precision=prec)
    bt.solve_dc(precision=prec, record_timing=2)
    ks = bt.kernel_stats()
    ty = "float>" if prec == _abi.NR_PREC_FP32_REFINE else "double>"
    # (instances the FP32 path hands to FP64 run the FP64 kernels, whose
    # launch mode is set separately)
    assert any(k.startswith("k_factor_persist") and k.endswith(ty) for k in ks), sorted(ks)
    assert not any(k.startswith("k_factor_large") and k.endswith(ty) for k in ks), sorted(ks)
    _same(ref, bt)
    # graph replay of the same sweeps
    bt.solve_dc(precision=prec, use_graph=1)
    _same(ref, bt)
    # and back to the per-level launches
    bt.set_persistent(False, precision=prec)
    bt.solve_dc(precision=prec, record_timing=2)
    assert not any(k.startswith("k_factor_persist") and k.endswith(ty) for k in bt.kernel_stats())
    _same(ref, bt)


@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
@pytest.mark.parametrize("name", ["rdmesh64", "logic2000"])
def test_fanout_equals_one_stream(ctx, name, prec):
    """Each level's launches on several streams (nr_batch_set_fanout), eager
    and replayed from hipGraphs, give the one-stream results bit for bit."""
    nl, B, kw = circuits()[name]
    f = netlists.mc_fields(nl, B, seed=45, **kw)
    ref, bt = _pair(ctx, nl, B, f)
    print(f"fanout {name} prec={prec} B={B}", flush=True)
    ref.solve_dc(precision=prec)
    bt.set_fanout(True, precision=prec)
    bt.solve_dc(precision=prec)
    _same(ref, bt)
    for _ in range(2):
        bt.solve_dc(precision=prec, use_graph=1)
        _same(ref, bt)
    bt.set_fanout(False, precision=prec)
    bt.solve_dc(precision=prec, use_graph=1)
    _same(ref, bt)


def test_persistent_transient(ctx):
    nl, B, kw = circuits()["logic2000"]
    B = 16
    f = netlists.mc_fields(nl, B, seed=43, **kw)
    ref, bt = _pair(ctx, nl, B, f)
    ref.solve_dc()
    V = ref.voltages()
    rt, tt = _pair(ctx, nl, B, f, transient=True)
    print("persistent transient logic2000", flush=True)
    tt.set_persistent(True)
    for b in (rt, tt):
        b.set_seed(V)
        b.transient(h=0.5e-9, steps=40, start=_abi.NR_TRAN_START_SEED, probes=[2, 3, 1000], use_graph=1)
    np.testing.assert_array_equal(tt.waveform(), rt.waveform())
    _same(rt, tt)


def test_autotune_times_the_persistent_kernels(ctx):
    nl, B, kw = circuits()["logic2000"]
    f = netlists.mc_fields(nl, B, seed=47, **kw)
    ref, bt = _pair(ctx, nl, B, f)
    print("persistent autotune logic2000", flush=True)
    ref.solve_dc()
    tune = bt.autotune(reps=2)
    assert tune["ms_levels"] > 0 and tune["ms_persistent"] > 0
    assert tune["ms_after"] == min(tune["ms_levels"], tune["ms_persistent"]) or tune["persistent"] == 0
    if not tune["persistent"]:
        assert tune["ms_graph"] > 0 and tune["ms_graph_fanout"] > 0
        assert tune["fanout"] == int(tune["ms_graph_fanout"] < 0.99 * tune["ms_graph"])
        # the second class map, tuned on 32 instances, serves the solve's tail
        assert tune["ms_few_before"] > 0 and tune["ms_few_after"] > 0
        assert tune["few_max"] in (0, 32, 64)
        assert (tune["few_max"] > 0) == (tune["ms_few_after"] < 0.99 * tune["ms_few_before"])
    print(f"few-instance map up to {tune['few_max']}: {tune['ms_few_before']:.2f} -> {tune['ms_few_after']:.2f} ms",
          flush=True)
    bt.solve_dc(use_graph=1)
    _same(ref, bt)
    bt.solve_dc()
    _same(ref, bt)


def test_bad_precision_is_refused(ctx):
    from karam_amd import nr
    nl = netlists.rd_mesh(8, seed=1)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), 4)
    with pytest.raises(nr.NrError):
        bt.set_persistent(True, precision=7)
