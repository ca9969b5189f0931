"""nr_solve_opts.polish on the GPU against the oracle: Newton iterations taken
after the last GMIN step converges, from the plain start and from a seed at
the last GMIN step; bit-identical on an exp-free circuit (an inverter chain),
iterations counted."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("polish", [0, 1, 2])
def test_polish_matches_the_oracle(polish):
    from karam_amd import nr
    from oracle import oracle as O
    ctx = nr.Context(0)
    nl = netlists.inverter_chain(16)
    B = 4
    f = netlists.mc_fields(nl, B, seed=9, sigma_vto=0.03, sigma_kp=0.03)
    bt = nr.Batch(ctx, nr.Topology(nl, ctx), B)
    bt.set_fields(f)
    To = O.Topology(nl)
    bt.solve_dc(polish=polish)
    V, st = bt.voltages(), bt.stats()
    Vo, so = O.solve_dc_batch(nl, To, f, B, polish=polish)
    np.testing.assert_array_equal(V, Vo)
    np.testing.assert_array_equal(st["iterations"], so["iterations"])
    seeds = V * (1.0 + 0.01 * np.cos(np.arange(V.size)).reshape(V.shape))
    bt.set_seed(seeds)
    bt.solve_dc(seeded_start_step=18, polish=polish)
    Vs, ss = bt.voltages(), bt.stats()
    Vso, sso = O.solve_dc_batch(nl, To, f, B, seeds=seeds, start_step=18, polish=polish)
    assert (ss["status"] == _abi.NR_STATUS_CONVERGED).all()
    np.testing.assert_array_equal(Vs, Vso)
    np.testing.assert_array_equal(ss["iterations"], sso["iterations"])
