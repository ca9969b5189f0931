"""Per-instance temperature corners (NR_FIELD_TEMP_K) on the GPU against the
oracle: a Monte-Carlo R-diode mesh whose instances sit at 250 K, 300.15 K
and 400 K, DC and a transient with grounded capacitors.

Diodes bring exp / log, so voltages are compared to 1e-9 rel + 1e-12 V and
iterations within +-1; an instance solved alone at its temperature equals its
row of the batch bit for bit, and the 300.15 K instances equal a batch
without the field."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu

TEMPS = np.array([250.0, 300.15, 400.0, 250.0, 400.0, 330.0])


def _close(a, b):
    return np.abs(a - b) <= 1e-9 * np.abs(b) + 1e-12


@pytest.fixture(scope="module")
def ctx():
    from karam_amd import nr
    return nr.Context(0)


def test_dc_corners_match_the_oracle(ctx):
    from karam_amd import nr
    from oracle import oracle as O
    nl = netlists.rd_mesh(24, seed=3, n_isrc=6)
    B = len(TEMPS)
    f = netlists.mc_fields(nl, B, seed=21)
    f[_abi.NR_FIELD_TEMP_K] = TEMPS
    T = nr.Topology(nl, ctx)
    bt = nr.Batch(ctx, T, B)
    bt.set_fields(f)
    bt.solve_dc()
    V, st = bt.voltages(), bt.stats()
    Vo, so = O.solve_dc_batch(nl, O.Topology(nl), f, B)
    print("corner iterations", st["iterations"], "oracle", so["iterations"], flush=True)
    assert (st["status"] == _abi.NR_STATUS_CONVERGED).all() and (st["kcl"] <= 1e-9).all()
    np.testing.assert_array_equal(st["status"], so["status"])
    assert np.abs(st["iterations"] - so["iterations"]).max() <= 1
    assert _close(V, Vo).all()
    # the corners move the answer: hot diodes drop more
    assert not np.allclose(V[0], V[2])
    # an instance alone at its temperature is its row of the batch
    for b in (0, 2):
        one = nr.Batch(ctx, T, 1)
        one.set_fields({k: np.asarray(v)[b:b + 1] for k, v in f.items()})
        one.solve_dc()
        np.testing.assert_array_equal(one.voltages()[0], V[b])
        one.close()
    # 300.15 K through the field equals the netlist's own ambient
    plain = nr.Batch(ctx, T, B)
    plain.set_fields({k: v for k, v in f.items() if k != _abi.NR_FIELD_TEMP_K})
    plain.solve_dc()
    np.testing.assert_array_equal(plain.voltages()[1], V[1])
    plain.close()
    # the FP32 path at the corners, within the voltage tolerance of FP64
    bt.solve_dc(precision=_abi.NR_PREC_FP32_REFINE)
    st32 = bt.stats()
    assert (st32["status"] == _abi.NR_STATUS_CONVERGED).all()
    assert _close(bt.voltages(), V).all()
    bt.close()


def test_transient_corners_match_the_oracle(ctx):
    from karam_amd import nr
    from karam_amd.netlists import _f64, _i32
    from oracle import oracle as O
    nl = netlists.rd_mesh(12, seed=5, n_isrc=4)
    n = nl.num_nodes
    nl.cap_n1, nl.cap_n2 = _i32(np.arange(1, n + 1)), _i32(np.zeros(n))
    nl.cap_farads = _f64(np.full(n, 2e-12))
    B, steps = len(TEMPS), 40
    f = netlists.mc_fields(nl, B, seed=8)
    f[_abi.NR_FIELD_TEMP_K] = TEMPS
    probes = [1, n // 2, n]
    Tt = nr.Topology(nl, ctx, transient=True)
    bt = nr.Batch(ctx, Tt, B)
    bt.set_fields(f)
    bt.transient(h=1e-9, steps=steps, probes=probes)
    W, st = bt.waveform(), bt.stats()
    bt.close()
    _, Wo, so = O.solve_transient_batch(nl, O.Topology(nl, transient=True), f, B, h=1e-9, steps=steps,
                                        probes=probes)
    np.testing.assert_array_equal(st["status"], so["status"])
    assert (st["status"] == _abi.NR_STATUS_CONVERGED).all()
    assert np.abs(st["iterations"] - so["iterations"]).max() <= steps // 10
    assert _close(W, Wo).all()
    assert not np.allclose(W[0, -1], W[2, -1])
