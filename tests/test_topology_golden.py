"""Symbolic analysis against the reference's own translation units.

The fixtures in tests/golden/topo_*.npz were produced by the reference's
tensorizeNetlist (tensors/physics_tensors.cpp:305-493), Min-Fill
(infrastructure/factor_graph.cpp:9-162, infrastructure/treewidth_analyzer.cpp:14-141),
MatrixConstructor::buildPattern (math/linalg.h:281-465) and symbolicFactorize
(math/sparse_lu.cpp:14-128), compiled from source.  Both the oracle and the
product library's host topology builder must reproduce every integer exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import TOPO_CASES, load_golden, netlist_from_golden
from karam_amd import _abi

E = _abi


def _builders():
    from karam_amd import nr
    from oracle import oracle as O
    return {"oracle": O.Topology, "product": nr.Topology}


@pytest.fixture(scope="module", params=TOPO_CASES)
def case(request):
    fx = load_golden("topo_" + request.param)
    return request.param, fx, netlist_from_golden(fx)


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_ordering_pattern_fill_bit_exact(case, who):
    name, fx, nl = case
    T = _builders()[who](nl)
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_ORDERING_RAW), fx["ref_ordering_raw"])
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_ROW_PTR), fx["ref_row_ptr"])
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_COL_IDX), fx["ref_col_idx"])
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_STAMP_SLOT), fx["ref_stamp_slot"])
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_PERM), fx["ref_perm"])
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_FILL_ROW_PTR), fx["ref_frow_ptr"])
    np.testing.assert_array_equal(T.export(E.NR_EXPORT_FILL_COL_IDX), fx["ref_fcol_idx"])


def test_product_info_matches_reference(case):
    from karam_amd import nr
    name, fx, nl = case
    info = nr.Topology(nl).info
    assert info["treewidth"] == int(fx["ref_treewidth"])
    assert info["fill_in"] == int(fx["ref_fill_in"])
    assert info["ordering_len"] == len(fx["ref_ordering_raw"])
    assert info["nnz_a"] == len(fx["ref_col_idx"])
    assert info["nnz_f"] == len(fx["ref_fcol_idx"])
    assert info["n_stamps"] == len(fx["ref_stamp_slot"])


@pytest.mark.parametrize("who", ["oracle", "product"])
def test_tensorize_device_order(case, who):
    """SoA order: devices bucketed by model name, netlist order inside a
    bucket (tensors/physics_tensors.cpp:481-490)."""
    name, fx, nl = case
    T = _builders()[who](nl)
    do = T.export(E.NR_EXPORT_DIODE_ORDER)
    mo = T.export(E.NR_EXPORT_MOS_ORDER)
    np.testing.assert_array_equal(nl.dio_anode[do], fx["ref_d_a"])
    np.testing.assert_array_equal(nl.dio_cathode[do], fx["ref_d_c"])
    np.testing.assert_array_equal(nl.dio_is[do], fx["ref_d_is"])
    np.testing.assert_array_equal(nl.mos_d[mo], fx["ref_m_d"])
    np.testing.assert_array_equal(nl.mos_g[mo], fx["ref_m_g"])
    np.testing.assert_array_equal(nl.mos_s[mo], fx["ref_m_s"])
    np.testing.assert_array_equal(nl.mos_b[mo], fx["ref_m_b"])
    kp, vto = nl.card_params()
    np.testing.assert_array_equal(kp[mo], fx["ref_m_kp"])
    np.testing.assert_array_equal(vto[mo], fx["ref_m_vth"])


def test_oracle_tensorize_parameters(case):
    """Every SoA parameter the oracle derives (diode Vt temperature scaling,
    MOSFET card defaults, PMOS flag, lambda) equals the reference's."""
    import ctypes as C
    from oracle import oracle as O
    name, fx, nl = case
    c, keep = nl.to_c()
    nd, nm = len(nl.dio_is), len(nl.mos_w)
    i32 = lambda k: np.zeros(max(k, 1), np.int32)  # noqa: E731
    f64 = lambda k: np.zeros(max(k, 1))  # noqa: E731
    d_a, d_c, d_is, d_n, d_vt = i32(nd), i32(nd), f64(nd), f64(nd), f64(nd)
    m_d, m_g, m_s, m_b = i32(nm), i32(nm), i32(nm), i32(nm)
    m_w, m_l, m_kp, m_vth, m_lam, m_pm = f64(nm), f64(nm), f64(nm), f64(nm), f64(nm), i32(nm)
    L = O.lib()
    L.orc_tensorize_export.argtypes = [C.POINTER(type(c))] + [C.c_void_p] * 15
    rc = L.orc_tensorize_export(C.byref(c), *[a.ctypes.data for a in (
        d_a, d_c, d_is, d_n, d_vt, m_d, m_g, m_s, m_b, m_w, m_l, m_kp, m_vth, m_lam, m_pm)])
    assert rc == 0
    for got, ref in ((d_a, "d_a"), (d_c, "d_c"), (d_is, "d_is"), (d_n, "d_n"), (d_vt, "d_vt"),
                     (m_d, "m_d"), (m_g, "m_g"), (m_s, "m_s"), (m_b, "m_b"), (m_w, "m_w"), (m_l, "m_l"),
                     (m_kp, "m_kp"), (m_vth, "m_vth"), (m_lam, "m_lam"), (m_pm, "m_pmos")):
        r = fx["ref_" + ref]
        np.testing.assert_array_equal(got[: len(r)], r, err_msg=ref)


def test_product_matches_oracle_on_full_mesh():
    """64x64 mesh (the Monte-Carlo workload): no reference fixture at this
    size, so the product's incremental Min-Fill is held to the oracle's
    literal restatement, itself pinned on the fixtures above."""
    from karam_amd import netlists, nr
    from oracle import oracle as O
    nl = netlists.rd_mesh(64, seed=1)
    T, P = O.Topology(nl), nr.Topology(nl)
    for what in (E.NR_EXPORT_ORDERING_RAW, E.NR_EXPORT_ROW_PTR, E.NR_EXPORT_COL_IDX, E.NR_EXPORT_STAMP_SLOT,
                 E.NR_EXPORT_PERM, E.NR_EXPORT_FILL_ROW_PTR, E.NR_EXPORT_FILL_COL_IDX):
        np.testing.assert_array_equal(P.export(what), T.export(what), err_msg=str(what))
    assert P.info["treewidth"] == T.treewidth
    assert P.info["nnz_f"] == 119948
