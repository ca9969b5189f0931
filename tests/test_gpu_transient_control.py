"""Transient companion methods, the reference's step control and the
waveform ring on the HIP path, against the oracle through the C ABI.

Bit-identical (voltages, times, iterations, step-control counters) where no
exp / log / sin is involved: inverter chains with PULSE inputs, FP64 and the
FP32 refinement path.  SINE sources bring sin and exp: voltages within
1e-9 relative + 1e-12 V, with the step decisions (the time grid) identical
in a run whose LTE stays far from its threshold."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists
from karam_amd.netlists import Netlist, _f64, _i32

pytestmark = pytest.mark.gpu
TRAP, GEAR1, GEAR2 = _abi.NR_INT_TRAP, _abi.NR_INT_GEAR1, _abi.NR_INT_GEAR2
ARB = _abi.NR_TRAN_ARBITER


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


@pytest.fixture(scope="module")
def O():
    from oracle import oracle
    return oracle


def gpu_run(ctx, nl, f, B, seed=None, **kw):
    from karam_amd import nr
    bt = nr.Batch(ctx, nr.Topology(nl, ctx, transient=True), B)
    bt.set_fields(f)
    if seed is not None:
        bt.set_seed(seed)
    res = bt.transient(**kw)
    return bt, res


def oracle_run(O, nl, f, B, **kw):
    kw = dict(kw)
    for a, b in (("wave_decimation", "decimation"),):
        if a in kw:
            kw[b] = kw.pop(a)
    if "v0" not in kw and kw.get("start", 0) == _abi.NR_TRAN_START_SEED:
        raise ValueError("seeded oracle run needs v0")
    return O.solve_transient_batch(nl, O.Topology(nl, transient=True), f, B, threads=4, full=True, **kw)


def assert_same(bt, ref, exact=True):
    Vo, Wo, so, To = ref
    st, ts = bt.stats(), bt.tran_stats()
    np.testing.assert_array_equal(st["status"], so["status"])
    np.testing.assert_array_equal(ts["accepted"], so["steps_done"])
    np.testing.assert_array_equal(ts["rejected"], so["rejected"])
    np.testing.assert_array_equal(ts["method"], so["method"])
    np.testing.assert_array_equal(ts["switched_at"], so["switched_at"])
    np.testing.assert_array_equal(ts["instability"], so["instability"])
    np.testing.assert_array_equal(ts["h_last"], so["h_last"])
    np.testing.assert_array_equal(bt.wave_recorded(), so["recorded"])
    np.testing.assert_array_equal(st["iterations"], so["iterations"])
    W, Tm, V = bt.waveform(), bt.waveform_times(), bt.voltages()
    if exact:
        np.testing.assert_array_equal(ts["t_end"], so["t_end"])
        np.testing.assert_array_equal(Tm, To)
        np.testing.assert_array_equal(W, Wo)
        np.testing.assert_array_equal(V, Vo)
        np.testing.assert_array_equal(st["kcl"], so["kcl"])
    else:
        np.testing.assert_allclose(ts["t_end"], so["t_end"], rtol=1e-15, atol=0)
        np.testing.assert_allclose(Tm, To, rtol=1e-15, atol=0)
        fin = np.isfinite(Wo)
        np.testing.assert_array_equal(np.isfinite(W), fin)
        assert np.all(np.abs(W[fin] - Wo[fin]) <= 1e-9 * np.abs(Wo[fin]) + 1e-12)


@pytest.mark.parametrize("method", [TRAP, GEAR2])
def test_fixed_step_methods_bit_exact(ctx, O, method):
    """inv16 Monte-Carlo, zero start, 120 fixed steps of 0.5 ns with the
    trapezoidal / Gear-2 companions."""
    nl = netlists.inverter_chain(16)
    B = 4
    f = netlists.mc_fields(nl, B, seed=31, sigma_vto=0.03, sigma_kp=0.03)
    probes = list(range(1, nl.num_nodes + 1))
    kw = dict(h=0.5e-9, steps=120, probes=probes, method=method)
    bt, res = gpu_run(ctx, nl, f, B, **kw)
    assert res["n_converged"] == B
    assert_same(bt, oracle_run(O, nl, f, B, **kw))


@pytest.mark.parametrize("start", ["zero", "dc"])
@pytest.mark.parametrize("prec", [_abi.NR_PREC_FP64, _abi.NR_PREC_FP32_REFINE])
def test_arbiter_bit_exact(ctx, O, start, prec):
    """inv12 Monte-Carlo under the reference's step control (trapezoidal at
    t = 0, t_stop = 60 ns, h = 1 ns): the same attempts, halvings, method
    switch, instability counts, time grid and waveforms as the oracle, from
    the zero start and from the DC operating point."""
    from karam_amd import nr
    nl = netlists.inverter_chain(12)
    B = 6
    f = netlists.mc_fields(nl, B, seed=12, sigma_vto=0.03, sigma_kp=0.03)
    probes = [2, 3, 8, nl.num_nodes]
    kw = dict(h=1e-9, method=TRAP, control=ARB, t_stop=60e-9, probes=probes, precision=prec)
    seed = None
    if start == "dc":
        D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
        D.set_fields(f)
        D.solve_dc(precision=prec)
        seed = D.voltages()
        kw["start"] = _abi.NR_TRAN_START_SEED
    bt, res = gpu_run(ctx, nl, f, B, seed=seed, **kw)
    okw = dict(kw)
    okw["precision"] = O.PREC_FP32_REFINE if prec == _abi.NR_PREC_FP32_REFINE else O.PREC_FP64
    if seed is not None:
        okw["v0"] = seed
    ref = oracle_run(O, nl, f, B, **okw)
    assert res["n_converged"] == B
    assert np.all(ref[2]["switched_at"] >= 0)       # the input edge forces the Gear-2 switch
    assert_same(bt, ref)


def test_arbiter_graph_replay_matches_eager(ctx):
    """Sweeps replayed from captured graphs give the eager results."""
    nl = netlists.inverter_chain(10)
    B = 5
    f = netlists.mc_fields(nl, B, seed=3, sigma_vto=0.03, sigma_kp=0.03)
    kw = dict(h=1e-9, method=TRAP, control=ARB, t_stop=40e-9, probes=[4, 9])
    a, _ = gpu_run(ctx, nl, f, B, **kw)
    g, _ = gpu_run(ctx, nl, f, B, use_graph=1, **kw)
    np.testing.assert_array_equal(a.waveform(), g.waveform())
    np.testing.assert_array_equal(a.waveform_times(), g.waveform_times())
    np.testing.assert_array_equal(a.stats()["iterations"], g.stats()["iterations"])
    for k, v in a.tran_stats().items():
        np.testing.assert_array_equal(v, g.tran_stats()[k])


def test_ring_wrap_and_decimation(ctx, O):
    """A 9-entry ring recording every 4th accepted step of an arbiter run:
    slots, times and values as the oracle's."""
    nl = netlists.inverter_chain(8)
    B = 3
    f = netlists.mc_fields(nl, B, seed=8, sigma_vto=0.03, sigma_kp=0.03)
    kw = dict(h=1e-9, method=TRAP, control=ARB, t_stop=80e-9, probes=[3, 6, 10], wave_capacity=9,
              wave_decimation=4)
    bt, _ = gpu_run(ctx, nl, f, B, **kw)
    ref = oracle_run(O, nl, f, B, **kw)
    assert np.all(ref[2]["recorded"] > 9)
    assert_same(bt, ref)


def test_newton_failure_path(ctx, O):
    """max_iter = 1: every attempt's Newton solve fails; four halvings end
    each instance at t = 0 with the oracle's status and counters."""
    nl = netlists.inverter_chain(4)
    B = 2
    f = netlists.mc_fields(nl, B, seed=1)
    kw = dict(h=1e-9, max_iter=1, method=TRAP, control=ARB, t_stop=10e-9, probes=[3])
    bt, res = gpu_run(ctx, nl, f, B, **kw)
    ref = oracle_run(O, nl, f, B, **kw)
    assert res["n_converged"] == 0
    assert np.all(bt.tran_stats()["rejected"] == 4)
    assert_same(bt, ref)


def sine_rc_mesh(side=8):
    """R-diode mesh with a capacitor at every node, driven by a cosine
    voltage source and by PULSE and SINE current sources."""
    nl = netlists.rd_mesh(side, seed=6)
    n = nl.num_nodes
    nl.cap_n1, nl.cap_n2 = _i32(np.arange(1, n + 1)), _i32(np.zeros(n))
    nl.cap_farads = _f64(np.full(n, 1e-12))
    nl.vsrc_pos, nl.vsrc_neg, nl.vsrc_volts = _i32([1]), _i32([0]), _f64([1.5])
    nl.vsrc_wave = _i32([_abi.NR_WAVE_SINE])
    nl.vsrc_wave_par = _f64([[0.5, 1.0, 20e6, 0.0, 0.0, 90.0, 0.0]])
    nl.isrc_pos, nl.isrc_neg, nl.isrc_amps = _i32([0, 0]), _i32([n // 2, n]), _f64([0.0, 1e-4])
    nl.isrc_wave = _i32([_abi.NR_WAVE_PULSE, _abi.NR_WAVE_SINE])
    nl.isrc_wave_par = _f64([[0.0, 2e-4, 5e-9, 1e-9, 1e-9, 10e-9, 30e-9],
                             [1e-4, 5e-5, 30e6, 2e-9, 1e7, 0.0, 0.0]])
    return nl


@pytest.mark.parametrize("method", [TRAP, GEAR1, GEAR2])
def test_sine_and_current_waveforms(ctx, O, method):
    """SINE V source, PULSE and damped SINE I sources on an 8x8 R-diode-C
    mesh, fixed steps from the DC operating point: voltages within
    tolerance of the oracle, same step outcomes."""
    from karam_amd import nr
    nl = sine_rc_mesh()
    B = 4
    f = netlists.mc_fields(nl, B, seed=17)
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.set_fields(f)
    D.solve_dc()
    seed = D.voltages()
    probes = [1, 2, 9, nl.num_nodes // 2, nl.num_nodes]
    kw = dict(h=0.25e-9, steps=80, probes=probes, method=method, start=_abi.NR_TRAN_START_SEED)
    bt, res = gpu_run(ctx, nl, f, B, seed=seed, **kw)
    ref = oracle_run(O, nl, f, B, v0=seed, **kw)
    assert res["n_converged"] == B
    Vo, Wo, so, To = ref
    np.testing.assert_array_equal(bt.stats()["status"], so["status"])
    np.testing.assert_array_equal(bt.tran_stats()["accepted"], so["steps_done"])
    np.testing.assert_allclose(bt.waveform_times(), To, rtol=0, atol=0)
    W = bt.waveform()
    assert np.all(np.abs(W - Wo) <= 1e-9 * np.abs(Wo) + 1e-12)


def test_cosine_rc_arbiter_grid(ctx, O):
    """A cosine-driven R-C pair from its operating point under the step
    control: every step stable (LTE far below 1), the same time grid as the
    oracle, voltages within tolerance."""
    from karam_amd import nr
    nl = Netlist(num_nodes=2, res_n1=_i32([1]), res_n2=_i32([2]), res_ohms=_f64([1e3]),
                 cap_n1=_i32([2]), cap_n2=_i32([0]), cap_farads=_f64([1e-12]),
                 vsrc_pos=_i32([1]), vsrc_neg=_i32([0]), vsrc_volts=_f64([1.0]),
                 vsrc_wave=_i32([_abi.NR_WAVE_SINE]), vsrc_wave_par=_f64([[0.0, 1.0, 1e6, 0, 0, 90.0, 0]]),
                 name="cos_rc")
    B = 3
    f = {_abi.NR_FIELD_RES_OHMS: np.array([[1e3], [2e3], [5e2]])}
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.set_fields(f)
    D.solve_dc()
    seed = D.voltages()
    kw = dict(h=1e-9, method=TRAP, control=ARB, t_stop=100e-9, probes=[1, 2], wave_capacity=128,
              start=_abi.NR_TRAN_START_SEED)
    bt, _ = gpu_run(ctx, nl, f, B, seed=seed, **kw)
    ref = oracle_run(O, nl, f, B, v0=seed, **kw)
    assert np.all(ref[2]["rejected"] == 0)
    assert_same(bt, ref, exact=False)
