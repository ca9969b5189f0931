"""SPICE text through the whole path: nr_spice_parse -> nr_topology_build ->
batched DC and transient on the GPU, against the oracle on the same parsed
netlist (bit-identical on the MOSFET chain; the diode circuit within
1e-9 rel + 1e-12 V)."""
from __future__ import annotations

import os

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "spice")


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def test_spice_inverter_chain_dc_and_transient(ctx):
    from karam_amd import nr, spice
    from oracle import oracle as O
    from test_spice import inverter_chain_spice
    nl, info = spice.parse(inverter_chain_spice(12))
    B = 4
    f = netlists.mc_fields(nl, B, seed=5, sigma_vto=0.03, sigma_kp=0.03)
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.set_fields(f)
    D.solve_dc()
    V = D.voltages()
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), f, B)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(D.stats()["iterations"], so["iterations"])
    tr = nr.Batch(ctx, nr.Topology(nl, ctx, transient=True), B)
    tr.set_fields(f)
    tr.set_seed(V)
    kw = dict(h=1e-9, method=_abi.NR_INT_TRAP, control=_abi.NR_TRAN_ARBITER, t_stop=50e-9, probes=[2, 5, 14])
    tr.transient(start=_abi.NR_TRAN_START_SEED, **kw)
    _, Wo, sto, To = O.solve_transient_batch(nl, O.Topology(nl, transient=True), f, B, start=1, v0=V, full=True,
                                             **kw)
    np.testing.assert_array_equal(tr.waveform(), Wo)
    np.testing.assert_array_equal(tr.waveform_times(), To)
    np.testing.assert_array_equal(tr.stats()["iterations"], sto["iterations"])


def test_spice_devices_netlist_dc(ctx):
    """devices.sp: cards, a diode clamp, a card-less PMOS-named model and a
    VSS-grounded device."""
    from karam_amd import nr, spice
    from oracle import oracle as O
    with open(os.path.join(HERE, "devices.sp")) as fh:
        nl, _ = spice.parse(fh.read())
    B = 3
    f = netlists.mc_fields(nl, B, seed=2)
    D = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    D.set_fields(f)
    res = D.solve_dc()
    V = D.voltages()
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), f, B)
    np.testing.assert_array_equal(D.stats()["status"], so["status"])
    assert res["n_converged"] == int((so["status"] == _abi.NR_STATUS_CONVERGED).sum())
    assert np.all(np.abs(V - Vo) <= 1e-9 * np.abs(Vo) + 1e-12)
