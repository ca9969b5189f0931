"""hipGraph replay of the Newton sweep (use_graph = 1): every replay equals
the eagerly launched sweep bit for bit.  The graphs are captured per grid
bucket (a power of two of instances) and reused across solves, so the tests
run several graph solves with eager solves and host synchronisation in
between, on batches whose active count shrinks through several buckets."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def _same(bt, ref):
    V, st = bt.voltages(), bt.stats()
    np.testing.assert_array_equal(V, ref[0])
    for k in ("iterations", "ref_iterations", "status", "gmin_steps_done", "kcl", "seeded"):
        np.testing.assert_array_equal(st[k], ref[1][k])


@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
def test_dc_graph_replays_equal_eager(ctx, prec):
    from karam_amd import nr
    nl = netlists.rd_mesh(16, seed=2)
    B = 300                                   # not a power of two; instances retire at different sweeps
    f = netlists.mc_fields(nl, B, seed=17)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    bt.set_fields(f)
    bt.solve_dc(precision=prec)
    ref = (bt.voltages(), bt.stats())
    for _ in range(3):
        r = bt.solve_dc(precision=prec, use_graph=1)
        assert r["n_converged"] == B
        _same(bt, ref)
        bt.solve_dc(precision=prec)           # eager work between the replays
        _same(bt, ref)


def test_seeded_fallback_graph_equals_eager(ctx):
    """Seeded start with a plain fallback for bad seeds, under graphs."""
    from karam_amd import nr
    nl = netlists.rd_mesh(16, seed=2)
    B = 70
    f = netlists.mc_fields(nl, B, seed=31)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    bt.set_fields(f)
    bt.solve_dc()
    seed = bt.voltages().copy()
    seed[40:] = 1e3
    bt.set_seed(seed)
    bt.solve_dc(seeded_start_step=18, max_iter=3)
    ref = (bt.voltages(), bt.stats())
    assert ref[1]["seeded"][:40].all() and not ref[1]["seeded"][40:].any()
    for _ in range(2):
        bt.solve_dc(seeded_start_step=18, max_iter=3, use_graph=1)
        _same(bt, ref)


def test_transient_graph_equals_eager(ctx):
    """Inverter chain, zero start, 120 steps: waveform and statistics."""
    from karam_amd import nr
    nl = netlists.inverter_chain(16)
    B = 6
    f = netlists.mc_fields(nl, B, seed=9, sigma_vto=0.03, sigma_kp=0.03)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx, transient=True), B)
    bt.set_fields(f)
    probes = list(range(1, nl.num_nodes + 1))
    bt.transient(h=1e-9, steps=120, probes=probes)
    W, ref = bt.waveform(), (bt.voltages(), bt.stats())
    for _ in range(2):
        bt.transient(h=1e-9, steps=120, probes=probes, use_graph=1)
        np.testing.assert_array_equal(bt.waveform(), W)
        _same(bt, ref)
