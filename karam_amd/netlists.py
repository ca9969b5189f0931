"""Synthetic netlists of the benchmark families, built programmatically.

Netlists are built the way the reference's own API builds them
(TensorBlock::addDiode / addMosfet, netlist/circuit.h:815-843), never through
the SPICE parser, whose diode and PULSE fields are left uninitialised.
Node ids are 1-based, 0 is ground.  Every generator is seeded (numpy PCG64)
so the oracle and the GPU library see identical inputs.

Families:
  inverter_chain  16-stage CMOS inverter chain (Level-1 NMOS/PMOS, 10 fF loads)
  rd_mesh         side x side resistor mesh, a diode to ground at every 4th
                  node, 8 current sources (the 64x64 Monte-Carlo workload)
  random_logic    levelised random NAND2/INV netlist with local fan-in
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _abi

VT_DEFAULT = 0.02585          # addDiode default thermal voltage (circuit.h:815)
AMBIENT_K = 300.15            # SimulationEnvironment default
WAVE_DC, WAVE_PULSE, WAVE_SINE = _abi.NR_WAVE_DC, _abi.NR_WAVE_PULSE, _abi.NR_WAVE_SINE


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


@dataclass
class ModelCard:
    name: str
    type: str | None            # "NMOS" / "PMOS" / "D"; None = name without a card
    kp: float | None = None
    vto: float | None = None
    lam: float | None = None


@dataclass
class Netlist:
    num_nodes: int
    ambient_temp_K: float = AMBIENT_K
    res_n1: np.ndarray = field(default_factory=lambda: _i32([]))
    res_n2: np.ndarray = field(default_factory=lambda: _i32([]))
    res_ohms: np.ndarray = field(default_factory=lambda: _f64([]))
    cap_n1: np.ndarray = field(default_factory=lambda: _i32([]))
    cap_n2: np.ndarray = field(default_factory=lambda: _i32([]))
    cap_farads: np.ndarray = field(default_factory=lambda: _f64([]))
    vsrc_pos: np.ndarray = field(default_factory=lambda: _i32([]))
    vsrc_neg: np.ndarray = field(default_factory=lambda: _i32([]))
    vsrc_volts: np.ndarray = field(default_factory=lambda: _f64([]))
    vsrc_wave: np.ndarray = field(default_factory=lambda: _i32([]))
    vsrc_wave_par: np.ndarray = field(default_factory=lambda: _f64(np.zeros((0, 7))))
    isrc_pos: np.ndarray = field(default_factory=lambda: _i32([]))
    isrc_neg: np.ndarray = field(default_factory=lambda: _i32([]))
    isrc_amps: np.ndarray = field(default_factory=lambda: _f64([]))
    isrc_wave: np.ndarray = field(default_factory=lambda: _i32([]))
    isrc_wave_par: np.ndarray = field(default_factory=lambda: _f64(np.zeros((0, 7))))
    dio_anode: np.ndarray = field(default_factory=lambda: _i32([]))
    dio_cathode: np.ndarray = field(default_factory=lambda: _i32([]))
    dio_is: np.ndarray = field(default_factory=lambda: _f64([]))
    dio_n: np.ndarray = field(default_factory=lambda: _f64([]))
    dio_vt: np.ndarray = field(default_factory=lambda: _f64([]))
    dio_model: np.ndarray | None = None
    mos_d: np.ndarray = field(default_factory=lambda: _i32([]))
    mos_g: np.ndarray = field(default_factory=lambda: _i32([]))
    mos_s: np.ndarray = field(default_factory=lambda: _i32([]))
    mos_b: np.ndarray = field(default_factory=lambda: _i32([]))
    mos_w: np.ndarray = field(default_factory=lambda: _f64([]))
    mos_l: np.ndarray = field(default_factory=lambda: _f64([]))
    mos_model: np.ndarray = field(default_factory=lambda: _i32([]))
    mos_kp: np.ndarray | None = None
    mos_vto: np.ndarray | None = None
    models: list = field(default_factory=list)
    name: str = "netlist"

    # -- counts --------------------------------------------------------------
    def field_count(self, fid: int) -> int:
        return {
            _abi.NR_FIELD_RES_OHMS: len(self.res_ohms),
            _abi.NR_FIELD_DIO_IS: len(self.dio_is),
            _abi.NR_FIELD_MOS_KP: len(self.mos_w),
            _abi.NR_FIELD_MOS_VTO: len(self.mos_w),
            _abi.NR_FIELD_ISRC_AMPS: len(self.isrc_amps),
            _abi.NR_FIELD_VSRC_VOLTS: len(self.vsrc_volts),
            _abi.NR_FIELD_CAP_FARADS: len(self.cap_farads),
            _abi.NR_FIELD_TEMP_K: 1,
        }[fid]

    def card_params(self):
        """Per-device (KP, VTO) that the tensorizer derives from the cards
        (tensors/physics_tensors.cpp:407-420), in netlist order."""
        m = len(self.mos_w)
        kp = np.full(m, 200e-6)
        vto = np.full(m, 0.7)
        any_card = any(c.type for c in self.models)
        for i in range(m):
            card = self.models[self.mos_model[i]] if self.mos_model[i] >= 0 else None
            name = card.name if card else ""
            if not any_card:
                continue
            real = next((c for c in self.models if c.type and c.name == name), None)
            if real is not None:
                pm = real.type == "PMOS"
                kp[i] = real.kp if real.kp is not None else (100e-6 if pm else 200e-6)
                vto[i] = real.vto if real.vto is not None else (-0.7 if pm else 0.7)
            elif "PMOS" in name or "pmos" in name:
                kp[i], vto[i] = 100e-6, -0.7
        if self.mos_kp is not None:
            kp = np.array(self.mos_kp, dtype=np.float64)
        if self.mos_vto is not None:
            vto = np.array(self.mos_vto, dtype=np.float64)
        return kp, vto

    # -- C view --------------------------------------------------------------
    def to_c(self):
        """Return (NrNetlist, keepalive).  keepalive must outlive the struct."""
        keep = []

        def p32(a):
            if a is None:
                return None
            a = _i32(a)
            keep.append(a)
            return a.ctypes.data_as(_abi._i32p)

        def p64(a):
            if a is None:
                return None
            a = _f64(a)
            keep.append(a)
            return a.ctypes.data_as(_abi._f64p)

        cards = (_abi.NrModelCard * max(1, len(self.models)))()
        for i, m in enumerate(self.models):
            nb = m.name.encode()
            tb = m.type.encode() if m.type else None
            keep.extend([nb, tb])
            cards[i] = _abi.NrModelCard(nb, tb, int(m.kp is not None), int(m.vto is not None),
                                        int(m.lam is not None), m.kp or 0.0, m.vto or 0.0,
                                        m.lam or 0.0)
        keep.append(cards)
        nl = _abi.NrNetlist()
        nl.num_nodes = int(self.num_nodes)
        nl.ambient_temp_K = float(self.ambient_temp_K)
        nl.n_res = len(self.res_ohms)
        nl.res_n1, nl.res_n2, nl.res_ohms = p32(self.res_n1), p32(self.res_n2), p64(self.res_ohms)
        nl.n_cap = len(self.cap_farads)
        nl.cap_n1, nl.cap_n2, nl.cap_farads = p32(self.cap_n1), p32(self.cap_n2), p64(self.cap_farads)
        nl.n_vsrc = len(self.vsrc_volts)
        nl.vsrc_pos, nl.vsrc_neg, nl.vsrc_volts = p32(self.vsrc_pos), p32(self.vsrc_neg), p64(self.vsrc_volts)
        if len(self.vsrc_wave) == nl.n_vsrc and nl.n_vsrc:
            nl.vsrc_wave = p32(self.vsrc_wave)
            nl.vsrc_wave_par = p64(np.asarray(self.vsrc_wave_par).reshape(-1))
        nl.n_isrc = len(self.isrc_amps)
        nl.isrc_pos, nl.isrc_neg, nl.isrc_amps = p32(self.isrc_pos), p32(self.isrc_neg), p64(self.isrc_amps)
        if len(self.isrc_wave) == nl.n_isrc and nl.n_isrc:
            nl.isrc_wave = p32(self.isrc_wave)
            nl.isrc_wave_par = p64(np.asarray(self.isrc_wave_par).reshape(-1))
        nl.n_diode = len(self.dio_is)
        nl.dio_anode, nl.dio_cathode = p32(self.dio_anode), p32(self.dio_cathode)
        nl.dio_is, nl.dio_n, nl.dio_vt = p64(self.dio_is), p64(self.dio_n), p64(self.dio_vt)
        nl.dio_model = p32(self.dio_model)
        nl.n_mos = len(self.mos_w)
        nl.mos_d, nl.mos_g, nl.mos_s, nl.mos_b = (p32(self.mos_d), p32(self.mos_g),
                                                  p32(self.mos_s), p32(self.mos_b))
        nl.mos_w, nl.mos_l, nl.mos_model = p64(self.mos_w), p64(self.mos_l), p32(self.mos_model)
        nl.mos_kp, nl.mos_vto = p64(self.mos_kp), p64(self.mos_vto)
        nl.n_models = len(self.models)
        nl.models = C.cast(cards, C.POINTER(_abi.NrModelCard))
        return nl, keep


# ---------------------------------------------------------------------------
# families
# ---------------------------------------------------------------------------
NMOS_CARD = ModelCard("NMOS", "NMOS", kp=110e-6, vto=0.7, lam=0.04)
PMOS_CARD = ModelCard("PMOS", "PMOS", kp=50e-6, vto=-0.7, lam=0.05)


def inverter_chain(stages: int = 16, vdd: float = 1.8, cload: float = 10e-15,
                   pulse=(0.0, 1.8, 20e-9, 1e-9, 1e-9, 1.0, 0.0)) -> Netlist:
    """CMOS inverter chain: VDD = node 1, IN = node 2, OUT_k = node 2 + k.
    NMOS W/L = 2/1, PMOS W/L = 4/1, a load capacitor per output; V_IN is a
    PULSE source whose DC value is its v1."""
    n = 2 + stages
    md, mg, ms, mb, mw, ml, mm = [], [], [], [], [], [], []
    for k in range(1, stages + 1):
        out = 2 + k
        gate = 2 if k == 1 else out - 1
        md += [out, out]; mg += [gate, gate]; ms += [0, 1]; mb += [0, 1]
        mw += [2e-6, 4e-6]; ml += [1e-6, 1e-6]; mm += [0, 1]
    return Netlist(
        num_nodes=n,
        cap_n1=_i32(range(3, n + 1)), cap_n2=_i32([0] * stages), cap_farads=_f64([cload] * stages),
        vsrc_pos=_i32([1, 2]), vsrc_neg=_i32([0, 0]), vsrc_volts=_f64([vdd, pulse[0]]),
        vsrc_wave=_i32([0, WAVE_PULSE]), vsrc_wave_par=_f64([[0] * 7, list(pulse)]),
        mos_d=_i32(md), mos_g=_i32(mg), mos_s=_i32(ms), mos_b=_i32(mb),
        mos_w=_f64(mw), mos_l=_f64(ml), mos_model=_i32(mm),
        models=[NMOS_CARD, PMOS_CARD], name=f"inv{stages}",
    )


def rd_mesh(side: int = 64, seed: int = 1, n_isrc: int = 8,
            r_lo: float = 1e3, r_hi: float = 10e3, i_lo: float = 1e-3, i_hi: float = 5e-3,
            diode_every: int = 4, Is: float = 1e-14) -> Netlist:
    """side x side resistor mesh, node id = 1 + side*r + c.  Resistors: all
    horizontal edges row-major, then all vertical edges; R ~ U(r_lo, r_hi).
    A diode (anode = node, cathode = ground) at every node with
    (id - 1) % diode_every == 0; n_isrc current sources (pos = ground,
    neg = node) inject U(i_lo, i_hi) into distinct seeded nodes."""
    rng = np.random.default_rng(seed)
    n = side * side
    ids = np.arange(1, n + 1).reshape(side, side)
    h1, h2 = ids[:, :-1].ravel(), ids[:, 1:].ravel()
    v1, v2 = ids[:-1, :].ravel(), ids[1:, :].ravel()
    n1 = np.concatenate([h1, v1]); n2 = np.concatenate([h2, v2])
    R = rng.uniform(r_lo, r_hi, size=n1.size)
    an = np.arange(1, n + 1)[(np.arange(1, n + 1) - 1) % diode_every == 0]
    isn = rng.choice(np.arange(1, n + 1), size=min(n_isrc, n), replace=False)
    I = rng.uniform(i_lo, i_hi, size=isn.size)
    return Netlist(
        num_nodes=n, res_n1=_i32(n1), res_n2=_i32(n2), res_ohms=_f64(R),
        isrc_pos=_i32(np.zeros(isn.size)), isrc_neg=_i32(isn), isrc_amps=_f64(I),
        dio_anode=_i32(an), dio_cathode=_i32(np.zeros(an.size)), dio_is=_f64(np.full(an.size, Is)),
        dio_n=_f64(np.ones(an.size)), dio_vt=_f64(np.full(an.size, VT_DEFAULT)),
        name=f"rdmesh{side}",
    )


def mc_fields(nl: Netlist, B: int, seed: int, sigma_r: float = 0.05, sigma_is: float = 0.05,
              sigma_vto: float = 0.0, sigma_kp: float = 0.0, sigma_isrc: float = 0.0,
              sigma_vsrc: float = 0.0) -> dict:
    """Per-instance Monte-Carlo parameter arrays [B, count] (netlist order):
    R and IS multiplied by max(0.05, 1 + sigma z); VTO += sigma_vto z,
    KP *= max(0.05, 1 + sigma_kp z); current and voltage source values (a
    corner's supply) multiplied by max(0.05, 1 + sigma z) when their sigma is
    set; z ~ N(0, 1), instance b drawn from seed + b so any instance can be
    regenerated on its own (the source draws come last, so they leave the
    other draws of an instance unchanged)."""
    out = {}
    nr, nd, nm = len(nl.res_ohms), len(nl.dio_is), len(nl.mos_w)
    R = np.empty((B, nr)); Isa = np.empty((B, nd))
    kp0, vto0 = nl.card_params()
    KP = np.empty((B, nm)); VT = np.empty((B, nm))
    ISRC = np.empty((B, len(nl.isrc_amps))); VSRC = np.empty((B, len(nl.vsrc_volts)))
    for b in range(B):
        rng = np.random.default_rng(seed + b)
        R[b] = nl.res_ohms * np.maximum(0.05, 1.0 + sigma_r * rng.standard_normal(nr))
        Isa[b] = nl.dio_is * np.maximum(0.05, 1.0 + sigma_is * rng.standard_normal(nd))
        if nm:
            VT[b] = vto0 + sigma_vto * rng.standard_normal(nm)
            KP[b] = kp0 * np.maximum(0.05, 1.0 + sigma_kp * rng.standard_normal(nm))
        if sigma_isrc:
            ISRC[b] = nl.isrc_amps * np.maximum(0.05, 1.0 + sigma_isrc * rng.standard_normal(len(nl.isrc_amps)))
        if sigma_vsrc:
            VSRC[b] = nl.vsrc_volts * np.maximum(0.05, 1.0 + sigma_vsrc * rng.standard_normal(len(nl.vsrc_volts)))
    if nr:
        out[_abi.NR_FIELD_RES_OHMS] = R
    if nd:
        out[_abi.NR_FIELD_DIO_IS] = Isa
    if nm and (sigma_vto or sigma_kp):
        out[_abi.NR_FIELD_MOS_KP] = KP
        out[_abi.NR_FIELD_MOS_VTO] = VT
    if sigma_isrc and len(nl.isrc_amps):
        out[_abi.NR_FIELD_ISRC_AMPS] = ISRC
    if sigma_vsrc and len(nl.vsrc_volts):
        out[_abi.NR_FIELD_VSRC_VOLTS] = VSRC
    return out


def instance(nl: Netlist, fields: dict, b: int) -> Netlist:
    """A copy of `nl` carrying instance b's values of `fields` (the netlist
    an instance of the batch solves)."""
    out = dataclasses.replace(nl)
    names = {_abi.NR_FIELD_RES_OHMS: "res_ohms", _abi.NR_FIELD_DIO_IS: "dio_is", _abi.NR_FIELD_MOS_KP: "mos_kp",
             _abi.NR_FIELD_MOS_VTO: "mos_vto", _abi.NR_FIELD_ISRC_AMPS: "isrc_amps",
             _abi.NR_FIELD_VSRC_VOLTS: "vsrc_volts", _abi.NR_FIELD_CAP_FARADS: "cap_farads"}
    for fid, arr in fields.items():
        if fid == _abi.NR_FIELD_TEMP_K:
            out.ambient_temp_K = float(np.asarray(arr, dtype=np.float64).reshape(-1)[b])
        else:
            setattr(out, names[fid], np.array(arr[b], dtype=np.float64))
    return out


def mc_features(nl: Netlist, fields: dict, n_features: int) -> np.ndarray:
    """Seeder input features of Monte-Carlo instances, [B, n_features] FP32:
    deviations from the template netlist (log ratios for R, IS, KP,
    capacitance and current-source values; differences for VTO and
    voltage-source values).  The independent sources' deviations come first,
    one feature each (a supply or an injected current moves every node);
    the devices' deviations, in netlist order and field order, are averaged
    over the remaining features in contiguous blocks."""
    kp0, vto0 = nl.card_params()
    nominal = {
        _abi.NR_FIELD_RES_OHMS: (nl.res_ohms, "log"), _abi.NR_FIELD_DIO_IS: (nl.dio_is, "log"),
        _abi.NR_FIELD_MOS_KP: (kp0, "log"), _abi.NR_FIELD_MOS_VTO: (vto0, "diff"),
        _abi.NR_FIELD_ISRC_AMPS: (nl.isrc_amps, "log"), _abi.NR_FIELD_VSRC_VOLTS: (nl.vsrc_volts, "diff"),
        _abi.NR_FIELD_CAP_FARADS: (nl.cap_farads, "log"),
    }
    src, dev = [], []
    for fid in sorted(fields):
        x = np.asarray(fields[fid], dtype=np.float64)
        base, kind = nominal[fid]
        if kind == "log":
            with np.errstate(divide="ignore", invalid="ignore"):
                d = np.log(np.abs(x) / np.abs(base))
            d[~np.isfinite(d)] = 0.0
        else:
            d = x - base
        (src if fid in (_abi.NR_FIELD_ISRC_AMPS, _abi.NR_FIELD_VSRC_VOLTS) else dev).append(d)
    if not src and not dev:
        raise ValueError("no perturbed fields")
    B = (src or dev)[0].shape[0]
    S = np.concatenate(src, axis=1) if src else np.zeros((B, 0))
    if S.shape[1] > n_features:
        raise ValueError("more perturbed sources than features")
    out = np.zeros((B, n_features), dtype=np.float32)
    out[:, :S.shape[1]] = S
    if dev:
        D = np.concatenate(dev, axis=1)
        nb = n_features - S.shape[1]
        for k, idx in enumerate(np.array_split(np.arange(D.shape[1]), nb)):
            if len(idx):
                out[:, S.shape[1] + k] = D[:, idx].mean(axis=1)
    return out


def random_logic(n_target: int = 20000, seed: int = 1, n_inputs: int = 32, window: int = 3, width: int = 32,
                 spread: int = 2, p_inv: float = 0.3, max_fanout: int = 4, vdd: float = 1.8,
                 cload: float = 10e-15, n_pulse: int = 8, p_nand3: float = 0.0) -> Netlist:
    """Levelised random CMOS logic of NAND2, NAND3 and INV gates (level-1 MOSFETs,
    the inverter-chain cards).  Node 1 is VDD; then n_inputs primary inputs
    driven by DC sources at 0 or VDD (seeded); every gate output has a
    `cload` capacitor to ground.  Gate inputs are drawn from the nets of the
    previous `window` levels (locality), each net driving at most
    `max_fanout` gates.  A NAND2 is two series NMOS (with an internal node)
    and two parallel PMOS; a NAND3 (probability p_nand3) three series NMOS
    (two internal nodes) and three parallel PMOS; an INV is one NMOS and one
    PMOS, so every gate brings two MOSFETs per node.  Gates are added
    until the node count reaches n_target.  Each level has `width` gate
    slots; slot p draws its inputs from slots p - spread .. p + spread of the
    previous `window` levels, which keeps the circuit banded (its
    elimination fronts grow with the width, not with the size).  The first
    n_pulse primary inputs are PULSE sources (0 -> VDD, seeded delay, width
    and period; DC value 0) so that a transient has switching activity."""
    rng = np.random.default_rng(seed)
    md, mg, ms, mb, mw, ml, mm = [], [], [], [], [], [], []
    cap_n = []
    node = 1                                   # VDD
    pis = list(range(2, 2 + n_inputs))
    node = 1 + n_inputs
    fan = {p: 0 for p in pis}
    levels = [pis]

    def nmos(d, g, s):
        md.append(d); mg.append(g); ms.append(s); mb.append(0); mw.append(2e-6); ml.append(1e-6); mm.append(0)

    def pmos(d, g, s):
        md.append(d); mg.append(g); ms.append(s); mb.append(1); mw.append(4e-6); ml.append(1e-6); mm.append(1)

    def near(lv, p):
        scale = len(lv) / float(width)
        c = int(p * scale)
        lo, hi = max(0, c - spread), min(len(lv), c + spread + 1)
        return lv[lo:hi]

    def pick(k, p):
        cand = [x for lv in levels[-window:] for x in near(lv, p)]
        pool = [x for x in cand if fan[x] < max_fanout]
        if len(pool) < k:
            pool = cand
        return list(rng.choice(pool, size=k, replace=False))

    while node < n_target:
        cur = []
        for p in range(width):
            if node >= n_target:
                break
            r = rng.random()
            if r < p_inv:
                (a,) = pick(1, p)
                out = node + 1; node = out
                nmos(out, a, 0); pmos(out, a, 1)
                fan[a] += 1
            elif r < p_inv + p_nand3:
                a, b, c = pick(3, p)
                out, m1, m2 = node + 1, node + 2, node + 3; node = m2
                nmos(out, a, m1); nmos(m1, b, m2); nmos(m2, c, 0)
                pmos(out, a, 1); pmos(out, b, 1); pmos(out, c, 1)
                fan[a] += 1; fan[b] += 1; fan[c] += 1
            else:
                a, b = pick(2, p)
                out, mid = node + 1, node + 2; node = mid
                nmos(out, a, mid); nmos(mid, b, 0)
                pmos(out, a, 1); pmos(out, b, 1)
                fan[a] += 1; fan[b] += 1
            fan[out] = 0
            cur.append(out)
            cap_n.append(out)
        levels.append(cur)
    pi_v = rng.integers(0, 2, n_inputs) * vdd
    n_pulse = min(n_pulse, n_inputs)
    pulse = np.zeros((1 + n_inputs, 7))
    wave = np.zeros(1 + n_inputs, dtype=np.int32)
    for k in range(n_pulse):
        per = rng.uniform(20e-9, 60e-9)
        pulse[1 + k] = [0.0, vdd, rng.uniform(1e-9, 10e-9), 0.5e-9, 0.5e-9, rng.uniform(0.3, 0.6) * per, per]
        wave[1 + k] = WAVE_PULSE
        pi_v[k] = 0.0
    return Netlist(
        num_nodes=node,
        cap_n1=_i32(cap_n), cap_n2=_i32(np.zeros(len(cap_n))), cap_farads=_f64(np.full(len(cap_n), cload)),
        vsrc_pos=_i32([1] + pis), vsrc_neg=_i32(np.zeros(1 + n_inputs)), vsrc_volts=_f64([vdd] + list(pi_v)),
        vsrc_wave=_i32(wave), vsrc_wave_par=_f64(pulse),
        mos_d=_i32(md), mos_g=_i32(mg), mos_s=_i32(ms), mos_b=_i32(mb),
        mos_w=_f64(mw), mos_l=_f64(ml), mos_model=_i32(mm),
        models=[NMOS_CARD, PMOS_CARD], name=f"logic{node}",
    )
