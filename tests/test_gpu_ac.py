"""AC small-signal analysis on the GPU (nr_batch_ac: one complex multifrontal
factorisation and solve of every instance per frequency) against the oracle's
restatement on the same operating points.  Bit-identical without diodes (the
complex kernels perform the oracle's operations in its order); with diodes
the conductances go through exp, and the results agree to 1e-9 relative."""
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


def run(ctx, nl, f, B, exact, input_node=1, probes=None, **kw):
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
    akw = dict(f_start=10.0, f_stop=10e9, n_points=13, input_node=input_node, amplitude=1.0)
    akw.update(kw)
    freqs, H, st = bt.ac(probes=probes, **akw)
    fo, Ho, so = O.solve_ac_batch(nl, O.Topology(nl, transient=True), f, B, V, probes=probes, threads=4, **akw)
    np.testing.assert_array_equal(freqs, fo)
    np.testing.assert_array_equal(st, so)
    assert np.all(st == _abi.NR_STATUS_CONVERGED)
    if exact:
        np.testing.assert_array_equal(H, Ho)
    else:
        np.testing.assert_allclose(H, Ho, rtol=1e-9, atol=1e-12 * np.abs(Ho).max())
    return H


def test_rc_ladder_bit_exact(ctx):
    rng = np.random.default_rng(1)
    n = 40
    nl = netlists.Netlist(num_nodes=n, res_n1=_i32(range(1, n)), res_n2=_i32(range(2, n + 1)),
                          res_ohms=_f64(rng.uniform(100, 5e3, n - 1)), cap_n1=_i32(range(2, n + 1)),
                          cap_n2=_i32([0] * (n - 1)), cap_farads=_f64(rng.uniform(1e-12, 1e-9, n - 1)),
                          vsrc_pos=_i32([1]), vsrc_neg=_i32([0]), vsrc_volts=_f64([1.0]), name="ladder")
    B = 8
    f = netlists.mc_fields(nl, B, seed=3)
    f[_abi.NR_FIELD_CAP_FARADS] = rng.uniform(1e-12, 1e-9, (B, n - 1))
    run(ctx, nl, f, B, exact=True)


def test_inverter_chain_bit_exact(ctx):
    """MOSFET small-signal (gm, gds of the instance's devices) at the chain's
    operating point, excited at its input."""
    nl = netlists.inverter_chain(10)
    B = 6
    f = netlists.mc_fields(nl, B, seed=4, sigma_vto=0.03, sigma_kp=0.03)
    run(ctx, nl, f, B, exact=True, input_node=2, probes=[2, 3, 7, 12])


@pytest.mark.parametrize("side", [12, 40])
def test_diode_mesh(ctx, side):
    """R-diode meshes with a capacitor at every node: small fronts and the
    large-front kernels on complex values (LDS and HBM panels)."""
    nl = with_caps(netlists.rd_mesh(side, seed=5))
    B = 16 if side == 12 else 4
    f = netlists.mc_fields(nl, B, seed=6)
    run(ctx, nl, f, B, exact=False, probes=[1, 2, side + 1, nl.num_nodes])


def test_clique_front_in_hbm(ctx):
    """A 210-node resistor clique (one front of order 210: its complex panel
    lives in HBM) with capacitors."""
    from test_gpu_bigfront import clique
    nl = with_caps(clique())
    B = 3
    f = netlists.mc_fields(nl, B, seed=8)
    run(ctx, nl, f, B, exact=False, probes=[1, 50, 210])
