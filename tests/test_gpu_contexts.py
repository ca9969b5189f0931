"""Contexts and topologies: a topology's device tables belong to the context
of its first batch, so a batch of it in a second context is refused (it
would otherwise re-upload the tables and leave the first context's batches
pointing at freed memory), while each context solves its own topology."""
from __future__ import annotations

import numpy as np
import pytest

from karam_amd import _abi, netlists

pytestmark = pytest.mark.gpu


def test_topology_belongs_to_the_context_of_its_first_batch():
    from karam_amd import nr
    c1, c2 = nr.Context(0), nr.Context(0)
    nl = netlists.rd_mesh(8, seed=1)
    f = netlists.mc_fields(nl, 2, seed=5)
    T = nr.Topology(nl, c1)
    b1 = nr.Batch(c1, T, 2)
    b1.set_fields(f)
    with pytest.raises(nr.NrError, match="another context"):
        nr.Batch(c2, T, 2)
    b2 = nr.Batch(c2, nr.Topology(nl, c2), 2)
    b2.set_fields(f)
    r1, r2 = b1.solve_dc(), b2.solve_dc()
    assert r1["n_converged"] == r2["n_converged"] == 2
    np.testing.assert_array_equal(b1.voltages(), b2.voltages())
    assert (b1.stats()["status"] == _abi.NR_STATUS_CONVERGED).all()
