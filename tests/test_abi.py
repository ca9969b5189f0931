"""The C-ABI boundary on the CPU: the library loads, exports exactly what
include/karam_nr.h declares, validates its inputs and reports errors as
status codes (no GPU work is done here)."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from conftest import ROOT
from karam_amd import _abi, netlists, nr


def _declared():
    txt = open(os.path.join(ROOT, "include", "karam_nr.h")).read()
    return sorted(set(re.findall(r"NR_API\s+[\w\s\*]*?\b(nr_\w+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert _declared() == sorted(nr.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(nr.LIB_PATH)
    for name in _declared():
        assert getattr(L, name, None) is not None, name
    assert nr.lib().nr_version().startswith(b"karam_nr")


def test_library_hides_internal_symbols():
    """-fvisibility=hidden: only nr_* entry points are dynamic symbols."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", nr.LIB_PATH], capture_output=True, text=True).stdout
    ours = [ln.split()[-1] for ln in out.splitlines() if " T " in ln]
    assert set(ours) >= set(_declared())
    assert not [s for s in ours if "build_topology" in s or "k_update" in s]


def test_ctx_create_status():
    h = C.c_void_p()
    rc = nr.lib().nr_ctx_create(C.byref(_abi.NrDeviceCfg(0, 0)), C.byref(h))
    assert rc in (_abi.NR_OK, _abi.NR_ERR_NO_DEVICE)
    if rc == _abi.NR_OK:
        nr.lib().nr_ctx_destroy(h)
    else:
        assert b"device" in nr.lib().nr_last_error(None)


def test_invalid_arguments_return_codes():
    L = nr.lib()
    h = C.c_void_p()
    assert L.nr_topology_build(None, None, C.byref(h)) == _abi.NR_ERR_INVALID
    assert L.nr_batch_alloc(None, None, 4, C.byref(h)) == _abi.NR_ERR_INVALID
    assert L.nr_batch_solve_dc(None, None, None) == _abi.NR_ERR_INVALID
    assert L.nr_topology_get_info(None, None) == _abi.NR_ERR_INVALID
    for bad in ("node", "ohms"):
        nl = netlists.rd_mesh(4, seed=1)
        if bad == "node":
            nl.res_n1 = nl.res_n1.copy(); nl.res_n1[0] = 99
        else:
            nl.res_ohms = nl.res_ohms.copy(); nl.res_ohms[3] = 0.0
        c, keep = nl.to_c()
        assert L.nr_topology_build(None, C.byref(c), C.byref(h)) == _abi.NR_ERR_INVALID
        assert b"resistor" in L.nr_last_error(None)


def test_export_size_query_and_capacity():
    nl = netlists.rd_mesh(8, seed=1)
    T = nr.Topology(nl)
    n = C.c_int64()
    assert nr.lib().nr_topology_export(T.h, _abi.NR_EXPORT_PERM, None, 0, C.byref(n)) == _abi.NR_OK
    assert n.value == 64
    small = np.zeros(10, np.int32)
    rc = nr.lib().nr_topology_export(T.h, _abi.NR_EXPORT_PERM, small.ctypes.data_as(C.POINTER(C.c_int32)), 10,
                                     C.byref(n))
    assert rc == _abi.NR_ERR_INVALID
    assert nr.lib().nr_topology_export(T.h, 77, None, 0, C.byref(n)) == _abi.NR_ERR_INVALID


def test_solve_opts_defaults_are_the_reference_plan():
    o = _abi.NrSolveOpts()
    nr.lib().nr_solve_opts_default(C.byref(o))
    assert (o.max_iter, o.gmin_steps, o.precision) == (50, 19, _abi.NR_PREC_FP64)


def test_mc_fields_reproducible_per_instance():
    nl = netlists.rd_mesh(8, seed=1)
    f = netlists.mc_fields(nl, 6, seed=100)
    g = netlists.mc_fields(nl, 3, seed=103)
    np.testing.assert_array_equal(f[_abi.NR_FIELD_RES_OHMS][3:], g[_abi.NR_FIELD_RES_OHMS])
    r = f[_abi.NR_FIELD_RES_OHMS] / nl.res_ohms
    assert 0.8 < r.min() and r.max() < 1.2 and abs(r.std() - 0.05) < 0.01


def test_product_does_not_reach_the_oracle():
    """The product package never imports, loads or names the oracle."""
    pkg = os.path.join(ROOT, "karam_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".inc")):
                txt = open(os.path.join(dirpath, f)).read()
                for pat in ("import oracle", "from oracle", "liboracle", "orc_", "oracle/", "karam_ref"):
                    assert pat not in txt, (f, pat)
