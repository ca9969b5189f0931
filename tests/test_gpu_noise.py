"""Noise analysis on the GPU (nr_batch_noise: per frequency one complex
factorisation of Y^T and one solve Y^T z = e_out for every instance, then the
sum over noise sources) against the oracle's restatements on the same
operating points:

  - the oracle's adjoint form performs the GPU's operations in its order:
    bit-identical PSDs without diodes (with diodes the shot noise and the
    conductances go through exp: 1e-9 relative);
  - the oracle's per-source form (the reference's one solve per source):
    1e-9 relative, the rounding of two different solve orders."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists
from karam_amd.netlists import _f64, _i32

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def with_caps(nl, c=2e-12):
    n = nl.num_nodes
    nl.cap_n1, nl.cap_n2 = _i32(np.arange(1, n + 1)), _i32(np.zeros(n))
    nl.cap_farads = _f64(np.full(n, c))
    return nl


def run(ctx, nl, f, B, exact, output_node, per_source=True, **kw):
    from karam_amd import nr
    from oracle import oracle as O
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.set_fields(f)
    D.solve_dc()
    V = D.voltages()
    T = nr.Topology(nl, ctx, transient=True)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    bt.set_seed(V)
    nkw = dict(f_start=10.0, f_stop=10e9, n_points=11, output_node=output_node, temp_K=300.15)
    nkw.update(kw)
    freqs, psd, vrms, st = bt.noise(**nkw)
    OT = O.Topology(nl, transient=True)
    fo, po, vo, so = O.solve_noise_batch(nl, OT, f, B, V, mode=O.NOISE_ADJOINT, threads=4, **nkw)
    np.testing.assert_array_equal(freqs, fo)
    np.testing.assert_array_equal(st, so)
    assert np.all(st == _abi.NR_STATUS_CONVERGED)
    assert np.all(psd > 0)
    if exact:
        np.testing.assert_array_equal(psd, po)
        np.testing.assert_allclose(vrms, vo, rtol=1e-15)
    else:
        np.testing.assert_allclose(psd, po, rtol=1e-9)
        np.testing.assert_allclose(vrms, vo, rtol=1e-9)
    if per_source:
        _, pp, vp, _ = O.solve_noise_batch(nl, OT, f, B, V, mode=O.NOISE_PER_SOURCE, threads=4, **nkw)
        np.testing.assert_allclose(psd, pp, rtol=1e-9)
        np.testing.assert_allclose(vrms, vp, rtol=1e-9)
    return psd, vrms


def test_rc_ladder_bit_exact(ctx):
    rng = np.random.default_rng(11)
    n = 40
    nl = netlists.Netlist(num_nodes=n, res_n1=_i32(range(1, n)), res_n2=_i32(range(2, n + 1)),
                          res_ohms=_f64(rng.uniform(100, 5e3, n - 1)), cap_n1=_i32(range(2, n + 1)),
                          cap_n2=_i32([0] * (n - 1)), cap_farads=_f64(rng.uniform(1e-12, 1e-9, n - 1)),
                          vsrc_pos=_i32([1]), vsrc_neg=_i32([0]), vsrc_volts=_f64([1.0]), name="ladder")
    B = 8
    f = netlists.mc_fields(nl, B, seed=13)
    f[_abi.NR_FIELD_CAP_FARADS] = rng.uniform(1e-12, 1e-9, (B, n - 1))
    f[_abi.NR_FIELD_RES_OHMS] = rng.uniform(100, 5e3, (B, n - 1))
    run(ctx, nl, f, B, exact=True, output_node=n)


def test_inverter_chain_bit_exact(ctx):
    """Channel thermal noise of every MOSFET at the chain's operating point
    (gm of the instance's devices); the output is the last stage."""
    nl = netlists.inverter_chain(10)
    B = 6
    f = netlists.mc_fields(nl, B, seed=14, sigma_vto=0.03, sigma_kp=0.03)
    run(ctx, nl, f, B, exact=True, output_node=nl.num_nodes)


@pytest.mark.parametrize("side", [12, 40])
def test_diode_mesh(ctx, side):
    """Shot noise of forward-biased diodes; small fronts and the large-front
    kernels factor Y^T on complex values (LDS and HBM panels)."""
    nl = with_caps(netlists.rd_mesh(side, seed=15))
    B = 16 if side == 12 else 4
    f = netlists.mc_fields(nl, B, seed=16)
    run(ctx, nl, f, B, exact=False, output_node=nl.num_nodes // 2, per_source=side == 12)


def test_clique_front_in_hbm(ctx):
    """A 210-node resistor clique (one front of order 210, its complex panel in
    HBM) with capacitors: 22,000 resistor sources summed per instance."""
    from test_gpu_bigfront import clique
    nl = with_caps(clique())
    B = 3
    f = netlists.mc_fields(nl, B, seed=18)
    run(ctx, nl, f, B, exact=True, output_node=100, per_source=False)


def test_output_node_zero_and_repeat(ctx):
    """Output node 0 gives zero noise; a second run on the same batch with
    other points reuses the buffers and returns the new points' results."""
    from karam_amd import nr
    nl = netlists.inverter_chain(4)
    B = 4
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.solve_dc()
    bt = nr.Batch(ctx, nr.Topology(nl, ctx, transient=True), B)
    bt.set_seed(D.voltages())
    _, p0, v0, _ = bt.noise(output_node=0, n_points=5)
    assert np.all(p0 == 0) and np.all(v0 == 0)
    fa, pa, _, _ = bt.noise(output_node=6, n_points=9, f_start=1e3, f_stop=1e6)
    fb, pb, _, _ = bt.noise(output_node=6, n_points=3, f_start=1e3, f_stop=1e6)
    assert pb.shape == (B, 3)
    np.testing.assert_array_equal(pb[:, 0], pa[:, 0])
    np.testing.assert_array_equal(pb[:, 2], pa[:, 8])
