"""The oracle's numerics against the reference's compiled kernels.

tests/golden/device_eval.npz holds outputs of the reference's generic device
kernels (tensors/physics_tensors.cpp:19-104), pnjlim / compute_vcrit
(physics/device_physics.h:132-179) and diode_current_clamped
(physics/device_physics.h:72-130) on grids that include every region
boundary; topo_*.npz hold a solveLU_Pivoted solution
(math/linalg.h:742-854) of each netlist's first Newton system.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import TOPO_CASES, load_golden, netlist_from_golden


@pytest.fixture(scope="module")
def dev():
    return load_golden("device_eval")


def test_diode_eval_bit_exact(oracle_lib, dev):
    O = oracle_lib
    vd, i_d, gd = O.eval_diodes(dev["dio_a"], dev["dio_c"], dev["dio_Is"], dev["dio_N"], dev["dio_Vt"], dev["dio_v"])
    np.testing.assert_array_equal(vd, dev["ref_dio_vd"])
    np.testing.assert_array_equal(i_d, dev["ref_dio_id"])
    np.testing.assert_array_equal(gd, dev["ref_dio_gd"])


def test_mosfet_eval_bit_exact(oracle_lib, dev):
    O = oracle_lib
    out = O.eval_mosfets(dev["mos_d"], dev["mos_g"], dev["mos_s"], dev["mos_W"], dev["mos_L"], dev["mos_Kp"],
                         dev["mos_Vth"], dev["mos_lam"], dev["mos_pm"], dev["mos_v"])
    for got, k in zip(out, ("vgs", "vds", "ids", "gm", "gds")):
        np.testing.assert_array_equal(got, dev["ref_mos_" + k], err_msg=k)


def test_pnjlim_vcrit_bit_exact(oracle_lib, dev):
    O = oracle_lib
    vc = np.array([O.vcrit(a, b) for a, b in zip(dev["pn_vt"], dev["pn_is"])])
    np.testing.assert_array_equal(vc, dev["ref_pn_vcrit"])
    lim = np.array([O.pnjlim(a, b, c, d) for a, b, c, d in
                    zip(dev["pn_vnew"], dev["pn_vold"], dev["pn_vt"], dev["ref_pn_vcrit"])])
    np.testing.assert_array_equal(lim, dev["ref_pn_lim"])


def test_diode_current_clamped_bit_exact(oracle_lib, dev):
    import ctypes as C
    L = oracle_lib.lib()
    L.orc_diode_i_clamped.argtypes = [C.c_double] * 4
    L.orc_diode_i_clamped.restype = C.c_double
    got = np.array([L.orc_diode_i_clamped(a, b, c, d) for a, b, c, d in
                    zip(dev["dc_vd"], dev["dc_is"], dev["dc_n"], dev["dc_vt"])])
    np.testing.assert_array_equal(got, dev["ref_dc_i"])


@pytest.mark.parametrize("name", TOPO_CASES)
def test_first_newton_system(oracle_lib, name):
    """Stamping + assembly at the seed: the oracle's CSR values, solved by
    the oracle's dense partial-pivot LU, give the reference's
    solveLU_Pivoted solution; the sparse and multifrontal LUs on the fill
    pattern agree with it to the solution tolerance."""
    O = oracle_lib
    fx = load_golden("topo_" + name)
    nl = netlist_from_golden(fx)
    T = O.Topology(nl)
    v0 = O.seed(nl)
    np.testing.assert_array_equal(v0, fx["lu_v0"])
    vals, av, rhs = O.stamp(nl, T, v0, 1e-3)
    np.testing.assert_allclose(av, fx["lu_a_vals"], rtol=1e-15, atol=0)
    np.testing.assert_allclose(rhs, fx["lu_rhs"], rtol=1e-15, atol=0)
    rc, x = O.lu_dense(nl.num_nodes, fx["ref_row_ptr"], fx["ref_col_idx"], fx["lu_a_vals"], fx["lu_rhs"])
    assert rc == int(fx["ref_lu_rc"])
    np.testing.assert_array_equal(x, fx["ref_lu_x"])
    tol = dict(rtol=1e-9, atol=1e-12)
    rc, xs = O.lu_sparse(T, fx["lu_a_vals"], fx["lu_rhs"])
    assert rc == 0
    np.testing.assert_allclose(xs, fx["ref_lu_x"], **tol)
    rc, xm = O.lu_multifrontal(T, fx["lu_a_vals"], fx["lu_rhs"])
    assert rc == 0
    np.testing.assert_allclose(xm, fx["ref_lu_x"], **tol)
