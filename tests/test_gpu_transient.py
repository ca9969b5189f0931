"""Transient parity of the HIP path with the oracle, through the C ABI.

Tolerances as in test_gpu_parity: voltages 1e-9 rel + 1e-12 V; iterations
within +-1 per step; bit-identical where no exp/log is involved (the
inverter chain)."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


@pytest.fixture(scope="module")
def O():
    from oracle import oracle
    return oracle


def test_inverter_chain_transient_bit_exact(ctx, O):
    """inv16 Monte-Carlo (VTO, KP), zero start, 200 steps of 1 ns."""
    from karam_amd import nr
    nl = netlists.inverter_chain(16)
    B, steps = 4, 200
    f = netlists.mc_fields(nl, B, seed=9, sigma_vto=0.03, sigma_kp=0.03)
    probes = list(range(1, nl.num_nodes + 1))
    T = nr.Topology(nl, ctx, transient=True)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    res = bt.transient(h=1e-9, steps=steps, probes=probes)
    W, V, st = bt.waveform(), bt.voltages(), bt.stats()
    Vo, Wo, so = O.solve_transient_batch(nl, O.Topology(nl, transient=True), f, B, h=1e-9, steps=steps,
                                         probes=probes)
    assert res["n_converged"] == B
    np.testing.assert_array_equal(W, Wo)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(st["iterations"], so["iterations"])
    np.testing.assert_array_equal(st["status"], so["status"])
    np.testing.assert_array_equal(st["gmin_steps_done"], so["steps_done"])
    np.testing.assert_array_equal(st["kcl"], so["kcl"])


def test_diode_mesh_with_capacitors(ctx, O):
    """16x16 R-diode mesh with a grounded capacitor at every node and a
    PULSE-driven source: diodes bring exp/log, so voltages are compared to
    the stated tolerance."""
    from karam_amd import nr
    from karam_amd.netlists import _f64, _i32
    nl = netlists.rd_mesh(16, seed=4)
    n = nl.num_nodes
    nl.cap_n1, nl.cap_n2 = _i32(np.arange(1, n + 1)), _i32(np.zeros(n))
    nl.cap_farads = _f64(np.full(n, 2e-12))
    nl.vsrc_pos, nl.vsrc_neg, nl.vsrc_volts = _i32([1]), _i32([0]), _f64([0.0])
    nl.vsrc_wave = _i32([_abi.NR_WAVE_PULSE])
    nl.vsrc_wave_par = _f64([[0.0, 2.0, 2e-9, 1e-9, 1e-9, 5e-9, 0.0]])
    B, steps, h = 6, 30, 0.5e-9
    f = netlists.mc_fields(nl, B, seed=21)
    f[_abi.NR_FIELD_CAP_FARADS] = np.random.default_rng(3).uniform(1e-12, 3e-12, (B, n))
    probes = [1, 2, 17, 100, n]
    T = nr.Topology(nl, ctx, transient=True)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    res = bt.transient(h=h, steps=steps, probes=probes)
    W, st = bt.waveform(), bt.stats()
    _, Wo, so = O.solve_transient_batch(nl, O.Topology(nl, transient=True), f, B, h=h, steps=steps,
                                        probes=probes)
    np.testing.assert_array_equal(st["status"], so["status"])
    assert res["n_converged"] == int((so["status"] == _abi.NR_STATUS_CONVERGED).sum()) == B
    assert np.all(np.abs(st["iterations"] - so["iterations"]) <= steps)
    err = np.abs(W - Wo) / (1e-9 * np.abs(Wo) + 1e-12)
    assert err.max() <= 1.0


@pytest.mark.parametrize("stages, steps", [(8, 40), (16, 200)])
def test_transient_from_dc_operating_point(ctx, O, stages, steps):
    """'spice' start: the DC operating point (DC topology) seeds the
    transient (transient topology) of the same instances; inv16 for 200
    steps of 1 ns is the single-instance configuration's run, here with
    Monte-Carlo instances."""
    from karam_amd import nr
    nl = netlists.inverter_chain(stages)
    B = 3
    print(f"transient from DC: inv{stages}, {steps} steps", flush=True)
    f = netlists.mc_fields(nl, B, seed=2, sigma_vto=0.02, sigma_kp=0.02)
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.set_fields(f)
    D.solve_dc()
    v0 = D.voltages()
    bt = nr.Batch(ctx, nr.Topology(nl, ctx, transient=True), B)
    bt.set_fields(f)
    bt.set_seed(v0)
    probes = list(range(1, nl.num_nodes + 1))
    bt.transient(h=1e-9, steps=steps, start=_abi.NR_TRAN_START_SEED, probes=probes)
    W, st = bt.waveform(), bt.stats()
    _, Wo, so = O.solve_transient_batch(nl, O.Topology(nl, transient=True), f, B, h=1e-9, steps=steps, start=1,
                                        v0=v0, probes=probes)
    np.testing.assert_array_equal(W, Wo)
    np.testing.assert_array_equal(W[:, 0, :], v0)
    np.testing.assert_array_equal(st["iterations"], so["iterations"])
    np.testing.assert_array_equal(st["status"], so["status"])


def test_kind_mismatch_is_an_error(ctx):
    from karam_amd import nr
    nl = netlists.inverter_chain(4)
    dc = nr.Batch(ctx, nr.Topology(nl, ctx), 1)
    tr = nr.Batch(ctx, nr.Topology(nl, ctx, transient=True), 1)
    with pytest.raises(nr.NrError):
        dc.transient(h=1e-9, steps=2)
    with pytest.raises(nr.NrError):
        tr.solve_dc()
