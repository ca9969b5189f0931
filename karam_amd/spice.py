"""SPICE text -> Netlist through the library's parser (nr_spice_parse, the
restated SPICEParser of netlist/spice_parser.cpp).  Host only: no GPU is
needed."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .netlists import ModelCard, Netlist


def _arr(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _name(fn, *args):
    buf = C.create_string_buffer(4096)
    rc = fn(*args, buf, len(buf))
    if rc != _abi.NR_OK:
        from .nr import _check
        _check(rc)
    return buf.value.decode()


def parse(text: str, vss_is_net: bool = False):
    """Parse SPICE text; returns (Netlist, info).  info holds the directives,
    counts, node names (index k-1 = node k), element names per kind, the V
    sources' AC magnitude / phase and the parser's warnings."""
    from .nr import _check, lib
    L = lib()
    raw = text.encode()
    h = C.c_void_p()
    _check(L.nr_spice_parse(raw, len(raw), _abi.NR_SPICE_VSS_IS_NET if vss_is_net else 0, C.byref(h)))
    try:
        c = _abi.NrNetlist()
        _check(L.nr_spice_get_netlist(h, C.byref(c)))
        inf = _abi.NrSpiceInfo()
        _check(L.nr_spice_get_info(h, C.byref(inf)))
        i32, f64 = np.int32, np.float64
        nV, nI = c.n_vsrc, c.n_isrc
        models = []
        for k in range(c.n_models):
            m = c.models[k]
            models.append(ModelCard(m.name.decode(), m.type.decode() if m.type else None,
                                    kp=m.kp if m.has_kp else None, vto=m.vto if m.has_vto else None,
                                    lam=m.lam if m.has_lambda else None))
        nl = Netlist(
            num_nodes=c.num_nodes, ambient_temp_K=c.ambient_temp_K,
            res_n1=_arr(c.res_n1, c.n_res, i32), res_n2=_arr(c.res_n2, c.n_res, i32),
            res_ohms=_arr(c.res_ohms, c.n_res, f64),
            cap_n1=_arr(c.cap_n1, c.n_cap, i32), cap_n2=_arr(c.cap_n2, c.n_cap, i32),
            cap_farads=_arr(c.cap_farads, c.n_cap, f64),
            vsrc_pos=_arr(c.vsrc_pos, nV, i32), vsrc_neg=_arr(c.vsrc_neg, nV, i32),
            vsrc_volts=_arr(c.vsrc_volts, nV, f64), vsrc_wave=_arr(c.vsrc_wave, nV, i32),
            vsrc_wave_par=_arr(c.vsrc_wave_par, 7 * nV, f64).reshape(nV, 7),
            isrc_pos=_arr(c.isrc_pos, nI, i32), isrc_neg=_arr(c.isrc_neg, nI, i32),
            isrc_amps=_arr(c.isrc_amps, nI, f64), isrc_wave=_arr(c.isrc_wave, nI, i32),
            isrc_wave_par=_arr(c.isrc_wave_par, 7 * nI, f64).reshape(nI, 7),
            dio_anode=_arr(c.dio_anode, c.n_diode, i32), dio_cathode=_arr(c.dio_cathode, c.n_diode, i32),
            dio_is=_arr(c.dio_is, c.n_diode, f64), dio_n=_arr(c.dio_n, c.n_diode, f64),
            dio_vt=_arr(c.dio_vt, c.n_diode, f64), dio_model=_arr(c.dio_model, c.n_diode, i32),
            mos_d=_arr(c.mos_d, c.n_mos, i32), mos_g=_arr(c.mos_g, c.n_mos, i32),
            mos_s=_arr(c.mos_s, c.n_mos, i32), mos_b=_arr(c.mos_b, c.n_mos, i32),
            mos_w=_arr(c.mos_w, c.n_mos, f64), mos_l=_arr(c.mos_l, c.n_mos, f64),
            mos_model=_arr(c.mos_model, c.n_mos, i32), models=models, name="spice",
        )
        counts = {"R": c.n_res, "C": c.n_cap, "V": nV, "I": nI, "D": c.n_diode, "M": c.n_mos}
        ac = []
        for k in range(nV):
            mag, ph = C.c_double(), C.c_double()
            _check(L.nr_spice_vsrc_ac(h, k, C.byref(mag), C.byref(ph)))
            ac.append((mag.value, ph.value))
        info = {f: getattr(inf, f) for f, _ in _abi.NrSpiceInfo._fields_ if f != "pad"}
        info["node_names"] = [_name(L.nr_spice_node_name, h, k) for k in range(1, c.num_nodes + 1)]
        info["element_names"] = {kind: [_name(L.nr_spice_element_name, h, kind.encode(), k) for k in range(n)]
                                 for kind, n in counts.items()}
        info["vsrc_ac"] = ac
        info["warning_text"] = [_name(L.nr_spice_warning, h, k) for k in range(inf.warnings)]
        return nl, info
    finally:
        L.nr_spice_free(h)
