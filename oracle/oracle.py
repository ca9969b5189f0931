"""ctypes binding of the C oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY — the checker.  Imported by tests/, by
__graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
product package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from karam_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class OrcConvStep(C.Structure):
    _fields_ = [("residual", C.c_double), ("max_dv", C.c_double), ("damping", C.c_double),
                ("gmin_step", C.c_int32), ("iteration", C.c_int32), ("worst_node", C.c_int32), ("pad", C.c_int32)]


class OrcDcOpts(C.Structure):
    _fields_ = [("max_iter", C.c_int32), ("gmin_steps", C.c_int32), ("lu_kind", C.c_int32),
                ("start_step", C.c_int32), ("use_seed", C.c_int32), ("precision", C.c_int32),
                ("trace_cap", C.c_int32), ("polish", C.c_int32), ("trace", C.POINTER(OrcConvStep))]


class OrcTranOpts(C.Structure):
    _fields_ = [("h", C.c_double), ("steps", C.c_int32), ("max_iter", C.c_int32), ("start", C.c_int32),
                ("lu_kind", C.c_int32), ("precision", C.c_int32), ("method", C.c_int32), ("control", C.c_int32),
                ("t_stop", C.c_double), ("wave_cap", C.c_int32), ("decimation", C.c_int32)]


class OrcTranOut(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("status", C.c_int32), ("steps_done", C.c_int32),
                ("rejected", C.c_int32), ("method", C.c_int32), ("switched_at", C.c_int32),
                ("instability", C.c_int32), ("taa_rejections", C.c_int32), ("recorded", C.c_int64),
                ("kcl", C.c_double), ("t_end", C.c_double), ("h_last", C.c_double),
                ("lu_fallbacks", C.c_int32), ("pad", C.c_int32)]


class OrcAcOpts(C.Structure):
    _fields_ = [("f_start", C.c_double), ("f_stop", C.c_double), ("n_points", C.c_int32),
                ("input_node", C.c_int32), ("amplitude", C.c_double)]


class OrcNoiseOpts(C.Structure):
    _fields_ = [("f_start", C.c_double), ("f_stop", C.c_double), ("n_points", C.c_int32),
                ("output_node", C.c_int32), ("temp_K", C.c_double), ("mode", C.c_int32), ("pad", C.c_int32)]


NOISE_PER_SOURCE, NOISE_ADJOINT = 0, 1


class OrcDcOut(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("ref_iterations", C.c_int32), ("status", C.c_int32),
                ("gmin_steps_done", C.c_int32), ("kcl", C.c_double),
                ("iters_per_step", C.c_int32 * 32),
                ("history", C.c_int64), ("failure", C.c_int32), ("stagnation", C.c_int32),
                ("divergence", C.c_int32), ("worst_node", C.c_int32), ("gmin_used", C.c_int32),
                ("lu_fallbacks", C.c_int32), ("rms", C.c_double), ("worst_v", C.c_double)]


FAIL_NONE, FAIL_SINGULAR, FAIL_DIVERGENCE, FAIL_LIMIT_CYCLE, FAIL_SLOW, FAIL_FLOATING, FAIL_OVERFLOW = range(7)


LU_DENSE = 0
LU_SPARSE = 1
LU_MULTIFRONTAL = 2
LU_AUTO = 3          # multifrontal, redone by the dense pivoting LU when its residual is large

PREC_FP64 = 0
PREC_FP32_REFINE = 1

INT_TRAP, INT_GEAR1, INT_GEAR2 = 0, 1, 2
TRAN_FIXED, TRAN_ARBITER = 0, 1


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_topology_build.argtypes = [C.POINTER(_abi.NrNetlist), C.POINTER(C.c_void_p)]
        L.orc_topology_build_ex.argtypes = [C.POINTER(_abi.NrNetlist), C.c_int32, C.POINTER(C.c_void_p)]
        L.orc_pulse_value.argtypes = [_f64p, C.c_double]
        L.orc_pulse_value.restype = C.c_double
        L.orc_sine_value.argtypes = [_f64p, C.c_double]
        L.orc_sine_value.restype = C.c_double
        L.orc_wave_capacity.argtypes = [C.POINTER(OrcTranOpts)]
        L.orc_ac_freqs.argtypes = [C.POINTER(OrcAcOpts), _f64p]
        L.orc_solve_ac_batch.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, C.c_int32, C.POINTER(_f64p), _f64p,
                                         C.POINTER(OrcAcOpts), C.c_int32, C.c_int32, _i32p, _f64p, _i32p]
        L.orc_solve_noise_batch.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, C.c_int32, C.POINTER(_f64p),
                                            _f64p, C.POINTER(OrcNoiseOpts), C.c_int32, _f64p, _f64p, _i32p]
        L.orc_arbiter_analyze.argtypes = [C.POINTER(_abi.NrNetlist), _f64p, _f64p, _f64p, C.c_double, _f64p, _f64p,
                                          _f64p, _f64p, C.POINTER(C.c_int32), _f64p]
        L.orc_wave_capacity.restype = C.c_int32
        L.orc_solve_transient_batch.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, C.c_int32,
                                                C.POINTER(_f64p), C.POINTER(OrcTranOpts), C.c_int32, _f64p,
                                                C.c_int32, _i32p, _f64p, _f64p, C.POINTER(OrcTranOut)]
        L.orc_topology_free.argtypes = [C.c_void_p]
        L.orc_topology_info.argtypes = [C.c_void_p, _i32p]
        L.orc_topology_export.argtypes = [C.c_void_p, C.c_int32, _i32p, C.c_int64]
        L.orc_topology_export.restype = C.c_int64
        L.orc_stamp.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, _f64p, C.c_double, _f64p, _f64p]
        L.orc_assemble.argtypes = [C.c_void_p, _f64p, _f64p]
        L.orc_lu_dense.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.orc_lu_sparse.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
        L.orc_seed.argtypes = [C.POINTER(_abi.NrNetlist), _f64p]
        L.orc_pnjlim.argtypes = [C.c_double] * 4
        L.orc_pnjlim.restype = C.c_double
        L.orc_vcrit.argtypes = [C.c_double, C.c_double]
        L.orc_vcrit.restype = C.c_double
        L.orc_apply_pn_limits.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, _f64p, _f64p]
        L.orc_kcl.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, _f64p, _f64p]
        L.orc_classify_history.argtypes = [_f64p, _f64p, C.c_int64, C.c_int32]
        L.orc_classify_history.restype = C.c_int32
        L.orc_kcl.restype = C.c_double
        L.orc_check_convergence.argtypes = [C.c_int32, _f64p, _f64p, C.c_double, C.c_double]
        L.orc_eval_diodes.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, C.c_int32,
                                      _f64p, _f64p, _f64p]
        L.orc_eval_mosfets.argtypes = [C.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                       _i32p, _f64p, C.c_int32, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.orc_solve_dc.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, C.POINTER(OrcDcOpts), _f64p,
                                   C.POINTER(OrcDcOut)]
        L.orc_solve_dc_batch.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, C.c_int32,
                                         C.POINTER(_f64p), C.POINTER(OrcDcOpts), C.c_int32, _f64p,
                                         C.POINTER(OrcDcOut)]
        _LIB = L
    return _LIB


def _p64(a):
    return a.ctypes.data_as(_f64p)


def _p32(a):
    return a.ctypes.data_as(_i32p)


class Topology:
    """Oracle topology: pattern, stamp map, Min-Fill ordering, fill pattern."""

    def __init__(self, nl, transient=False, ordering=None):
        """ordering: the raw Min-Fill ordering of this netlist from an earlier
        build (NR_EXPORT_ORDERING_RAW), which skips the literal Min-Fill."""
        self.nl = nl
        self._c, self._keep = nl.to_c()
        h = C.c_void_p()
        if ordering is None:
            rc = lib().orc_topology_build_ex(C.byref(self._c), 1 if transient else 0, C.byref(h))
        else:
            o = np.ascontiguousarray(ordering, np.int32)
            lib().orc_topology_build_ordered.argtypes = [C.POINTER(_abi.NrNetlist), C.c_int32, _i32p, C.c_int32,
                                                         C.POINTER(C.c_void_p)]
            rc = lib().orc_topology_build_ordered(C.byref(self._c), 1 if transient else 0, _p32(o), len(o),
                                                  C.byref(h))
        if rc:
            raise RuntimeError(f"orc_topology_build failed: {rc}")
        self.h = h
        info = np.zeros(8, dtype=np.int32)
        lib().orc_topology_info(h, _p32(info))
        (self.n, self.nnz_a, self.nnz_f, self.n_stamps, self.treewidth,
         self.ordering_len, self.fill_in, self.n_vars) = [int(x) for x in info]

    def export(self, what):
        cap = max(self.n + 1, self.nnz_a, self.nnz_f, self.n_stamps, self.ordering_len,
                  len(self.nl.dio_is), len(self.nl.mos_w)) + 1
        buf = np.zeros(cap, dtype=np.int32)
        k = lib().orc_topology_export(self.h, what, _p32(buf), cap)
        if k < 0:
            raise RuntimeError("export failed")
        return buf[:k].copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_topology_free(self.h)
            self.h = None


def dc_opts(max_iter=50, gmin_steps=19, lu=LU_AUTO, start_step=0, use_seed=0, precision=PREC_FP64,
            trace_cap=0, trace=None, polish=0):
    return OrcDcOpts(max_iter, gmin_steps, lu, start_step, use_seed, precision, trace_cap, polish, trace)


def classify_history(residual, max_dv, singular=False):
    """The oracle's incremental failure classification of a history."""
    r = np.ascontiguousarray(residual, np.float64)
    d = np.ascontiguousarray(max_dv, np.float64)
    return lib().orc_classify_history(_p64(r), _p64(d), C.c_int64(len(r)), int(singular))


def _trace_dict(tr):
    """ConvergenceStep records as arrays (numpy view of the ctypes array)."""
    a = np.ctypeslib.as_array(tr).view(np.dtype([("residual", "<f8"), ("max_dv", "<f8"), ("damping", "<f8"),
                                                 ("gmin_step", "<i4"), ("iteration", "<i4"),
                                                 ("worst_node", "<i4"), ("pad", "<i4")]))
    return {k: np.array(a[k]) for k in ("residual", "max_dv", "damping", "gmin_step", "iteration", "worst_node")}


def solve_dc(nl, topo: Topology | None = None, lu=LU_AUTO, max_iter=50, gmin_steps=19,
             seed_v=None, start_step=0, precision=PREC_FP64):
    c, keep = nl.to_c()
    v = np.zeros(nl.num_nodes) if seed_v is None else np.array(seed_v, dtype=np.float64)
    out = OrcDcOut()
    o = dc_opts(max_iter, gmin_steps, lu, start_step, 0 if seed_v is None else 1, precision)
    lib().orc_solve_dc(C.byref(c), topo.h if topo else None, C.byref(o), _p64(v), C.byref(out))
    return v, out


def solve_dc_batch(nl, topo: Topology, fields: dict, B: int, lu=LU_AUTO, threads=0,
                   max_iter=50, gmin_steps=19, precision=PREC_FP64, seeds=None, start_step=0, trace_cap=0,
                   polish=0):
    """Solve B Monte-Carlo instances; fields[fid] is a [B, count] array;
    seeds (optional, [B, n]) start every instance at GMIN step start_step.
    trace_cap > 0 adds stats["trace"]: the last trace_cap ConvergenceStep
    records of every instance ([B, trace_cap] arrays, record k at k %
    trace_cap)."""
    c, keep = nl.to_c()
    arrs = (_f64p * _abi.N_FIELDS)()
    hold = []
    for fid in range(_abi.N_FIELDS):
        if fid in fields:
            a = np.ascontiguousarray(fields[fid][:B], dtype=np.float64)
            hold.append(a)
            arrs[fid] = _p64(a)
        else:
            arrs[fid] = None
    v = np.zeros((B, nl.num_nodes)) if seeds is None else np.array(seeds, dtype=np.float64).reshape(B, -1)
    outs = (OrcDcOut * B)()
    tr = (OrcConvStep * (B * trace_cap))() if trace_cap > 0 else None
    o = dc_opts(max_iter, gmin_steps, lu, start_step if seeds is not None else 0, 0 if seeds is None else 1,
                precision, trace_cap, C.cast(tr, C.POINTER(OrcConvStep)) if tr is not None else None, polish)
    lib().orc_solve_dc_batch(C.byref(c), topo.h, B, arrs, C.byref(o), threads, _p64(v), outs)
    stats = {
        "iterations": np.array([x.iterations for x in outs]),
        "ref_iterations": np.array([x.ref_iterations for x in outs]),
        "status": np.array([x.status for x in outs]),
        "kcl": np.array([x.kcl for x in outs]),
        "gmin_steps_done": np.array([x.gmin_steps_done for x in outs]),
        "history": np.array([x.history for x in outs]),
        "failure": np.array([x.failure for x in outs]),
        "stagnation": np.array([x.stagnation for x in outs]),
        "divergence": np.array([x.divergence for x in outs]),
        "worst_node": np.array([x.worst_node for x in outs]),
        "worst_v": np.array([x.worst_v for x in outs]),
        "rms": np.array([x.rms for x in outs]),
        "gmin_used": np.array([x.gmin_used for x in outs]),
        "lu_fallbacks": np.array([x.lu_fallbacks for x in outs]),
    }
    if tr is not None:
        stats["trace"] = {k: x.reshape(B, trace_cap) for k, x in _trace_dict(tr).items()}
    return v, stats


def pulse_value(p, t):
    p = np.ascontiguousarray(p, np.float64)
    return lib().orc_pulse_value(_p64(p), t)


def sine_value(p, t):
    p = np.ascontiguousarray(p, np.float64)
    return lib().orc_sine_value(_p64(p), t)


def arbiter_analyze(nl, cur, hv0, hv1, dt, prev, prev2, prev_energy, prev_dt, instab):
    """IntegratorArbiter::analyzeStability on a given state; returns
    (suggestion, lte, state) with state = dict(prev, prev2, prev_energy,
    prev_dt, instab) after the call."""
    c, keep = nl.to_c()
    cur, hv0, hv1 = (np.ascontiguousarray(x, np.float64) for x in (cur, hv0, hv1))
    prev, prev2 = np.array(prev, np.float64), np.array(prev2, np.float64)
    pe, pdt, ins, lte = C.c_double(prev_energy), C.c_double(prev_dt), C.c_int32(instab), C.c_double(0.0)
    r = lib().orc_arbiter_analyze(C.byref(c), _p64(cur), _p64(hv0) if len(hv0) else None,
                                  _p64(hv1) if len(hv1) else None, dt, _p64(prev), _p64(prev2), C.byref(pe),
                                  C.byref(pdt), C.byref(ins), C.byref(lte))
    return r, lte.value, {"prev": prev, "prev2": prev2, "prev_energy": pe.value, "prev_dt": pdt.value,
                          "instab": ins.value}


def solve_transient_batch(nl, topo: Topology, fields: dict, B: int, h=1e-9, steps=200, max_iter=50, start=0,
                          v0=None, probes=None, lu=LU_AUTO, threads=0, precision=PREC_FP64,
                          method=INT_GEAR1, control=TRAN_FIXED, t_stop=0.0, wave_capacity=0, decimation=1,
                          full=False):
    """Transient of B instances; returns (v_final [B, n], wave [B, cap, n_probe],
    stats) and, with full=True, also the ring's times [B, cap].  The ring
    holds entry k in slot k mod cap (entry 0 = the start state)."""
    c, keep = nl.to_c()
    arrs = (_f64p * _abi.N_FIELDS)()
    hold = []
    for fid in range(_abi.N_FIELDS):
        if fid in fields:
            a = np.ascontiguousarray(fields[fid][:B], dtype=np.float64)
            hold.append(a)
            arrs[fid] = _p64(a)
        else:
            arrs[fid] = None
    v = np.zeros((B, nl.num_nodes)) if v0 is None else np.array(v0, dtype=np.float64).reshape(B, nl.num_nodes)
    pr = np.ascontiguousarray(probes if probes is not None else [], np.int32)
    o = OrcTranOpts(h, steps, max_iter, start, lu, precision, method, control, t_stop, wave_capacity, decimation)
    cap = lib().orc_wave_capacity(C.byref(o))
    wave = np.zeros((B, cap, len(pr)))
    times = np.zeros((B, cap))
    outs = (OrcTranOut * B)()
    rc = lib().orc_solve_transient_batch(C.byref(c), topo.h, B, arrs, C.byref(o), threads, _p64(v), len(pr),
                                         _p32(pr) if len(pr) else None, _p64(times), _p64(wave), outs)
    if rc != 0:
        raise ValueError(f"orc_solve_transient_batch: invalid arguments (status {rc})")
    stats = {k: np.array([getattr(x, k) for x in outs])
             for k in ("iterations", "status", "steps_done", "rejected", "method", "switched_at", "instability",
                       "taa_rejections", "recorded", "kcl", "t_end", "h_last", "lu_fallbacks")}
    if full:
        return v, wave, stats, times
    return v, wave, stats


def solve_ac_batch(nl, topo: Topology, fields: dict, B: int, v_op, f_start=1.0, f_stop=1e9, n_points=50,
                   input_node=0, amplitude=1.0, probes=None, threads=0):
    """AC sweep of B instances at their operating points v_op [B, n];
    returns (freqs [F], H [B, F, n_probe] complex, status [B])."""
    c, keep = nl.to_c()
    arrs = (_f64p * _abi.N_FIELDS)()
    hold = []
    for fid in range(_abi.N_FIELDS):
        if fid in fields:
            a = np.ascontiguousarray(fields[fid][:B], dtype=np.float64)
            hold.append(a)
            arrs[fid] = _p64(a)
        else:
            arrs[fid] = None
    v = np.ascontiguousarray(v_op, dtype=np.float64).reshape(B, nl.num_nodes)
    pr = np.ascontiguousarray(probes if probes is not None else np.arange(1, nl.num_nodes + 1), np.int32)
    o = OrcAcOpts(f_start, f_stop, n_points, input_node, amplitude)
    freqs = np.zeros(n_points)
    lib().orc_ac_freqs(C.byref(o), _p64(freqs))
    out = np.zeros((B, n_points, len(pr), 2))
    status = np.zeros(B, dtype=np.int32)
    rc = lib().orc_solve_ac_batch(C.byref(c), topo.h, B, arrs, _p64(v), C.byref(o), threads, len(pr), _p32(pr),
                                  _p64(out), _p32(status))
    if rc != 0:
        raise ValueError(f"orc_solve_ac_batch: invalid arguments (status {rc})")
    return freqs, out[..., 0] + 1j * out[..., 1], status


def solve_noise_batch(nl, topo: Topology, fields: dict, B: int, v_op, f_start=1.0, f_stop=1e9, n_points=50,
                      output_node=1, temp_K=300.15, mode=NOISE_PER_SOURCE, threads=0):
    """Noise analysis of B instances at their operating points v_op [B, n]:
    (freqs [F], output noise PSD [B, F] in V^2/Hz, integrated RMS noise [B],
    status [B]).  mode NOISE_PER_SOURCE is the reference's one solve per noise
    source, NOISE_ADJOINT one transposed solve per frequency."""
    c, keep = nl.to_c()
    arrs = (_f64p * _abi.N_FIELDS)()
    hold = []
    for fid in range(_abi.N_FIELDS):
        if fid in fields:
            a = np.ascontiguousarray(fields[fid][:B], dtype=np.float64)
            hold.append(a)
            arrs[fid] = _p64(a)
        else:
            arrs[fid] = None
    v = np.ascontiguousarray(v_op, dtype=np.float64).reshape(B, nl.num_nodes)
    o = OrcNoiseOpts(f_start, f_stop, n_points, output_node, temp_K, mode, 0)
    freqs = np.zeros(n_points)
    lib().orc_ac_freqs(C.byref(OrcAcOpts(f_start, f_stop, n_points, 0, 1.0)), _p64(freqs))
    psd = np.zeros((B, n_points))
    vrms = np.zeros(B)
    status = np.zeros(B, dtype=np.int32)
    rc = lib().orc_solve_noise_batch(C.byref(c), topo.h, B, arrs, _p64(v), C.byref(o), threads, _p64(psd),
                                     _p64(vrms), _p32(status))
    if rc != 0:
        raise ValueError(f"orc_solve_noise_batch: invalid arguments (status {rc})")
    return freqs, psd, vrms, status


def eval_diodes(a, c, Is, N, Vt, v):
    a, c = np.ascontiguousarray(a, np.int32), np.ascontiguousarray(c, np.int32)
    Is, N, Vt, v = (np.ascontiguousarray(x, np.float64) for x in (Is, N, Vt, v))
    m = len(a)
    vd, i_d, gd = np.zeros(m), np.zeros(m), np.zeros(m)
    lib().orc_eval_diodes(m, _p32(a), _p32(c), _p64(Is), _p64(N), _p64(Vt), _p64(v), len(v),
                          _p64(vd), _p64(i_d), _p64(gd))
    return vd, i_d, gd


def eval_mosfets(d, g, s, W, L, Kp, Vth, lam, pmos, v):
    d, g, s, pmos = (np.ascontiguousarray(x, np.int32) for x in (d, g, s, pmos))
    W, L, Kp, Vth, lam, v = (np.ascontiguousarray(x, np.float64) for x in (W, L, Kp, Vth, lam, v))
    m = len(d)
    outs = [np.zeros(m) for _ in range(5)]
    lib().orc_eval_mosfets(m, _p32(d), _p32(g), _p32(s), _p64(W), _p64(L), _p64(Kp), _p64(Vth),
                           _p64(lam), _p32(pmos), _p64(v), len(v), *[_p64(o) for o in outs])
    return tuple(outs)


def kcl(nl, v, per_node=False):
    c, keep = nl.to_c()
    v = np.ascontiguousarray(v, np.float64)
    res = np.zeros(nl.num_nodes)
    m = lib().orc_kcl(C.byref(c), None, _p64(v), _p64(res))
    return (m, res) if per_node else m


def stamp(nl, topo: Topology, v, gmin):
    c, keep = nl.to_c()
    v = np.ascontiguousarray(v, np.float64)
    vals = np.zeros(topo.n_stamps)
    rhs = np.zeros(nl.num_nodes)
    rc = lib().orc_stamp(C.byref(c), topo.h, _p64(v), gmin, _p64(vals), _p64(rhs))
    if rc:
        raise RuntimeError("orc_stamp failed")
    av = np.zeros(topo.nnz_a)
    lib().orc_assemble(topo.h, _p64(vals), _p64(av))
    return vals, av, rhs


def node_conductance(nl, topo: Topology, v, gmin=1e-12):
    """Diagonal of the DC Jacobian at v (S per node): how firmly the circuit
    holds each node.  A voltage difference d at node i is a current of about
    G_ii |d| there, which compares with the KCL acceptance of 1e-9 A."""
    _, av, _ = stamp(nl, topo, v, gmin)
    rp = topo.export(_abi.NR_EXPORT_ROW_PTR)
    ci = topo.export(_abi.NR_EXPORT_COL_IDX)
    rows = np.repeat(np.arange(topo.n), np.diff(rp))
    d = np.zeros(topo.n)
    on = ci == rows
    d[rows[on]] = av[on]
    return d


def lu_dense(n, rp, ci, vals, rhs):
    x = np.zeros(n)
    sp = C.c_double()
    rc = lib().orc_lu_dense(n, _p32(np.ascontiguousarray(rp, np.int32)), _p32(np.ascontiguousarray(ci, np.int32)),
                            _p64(np.ascontiguousarray(vals, np.float64)),
                            _p64(np.ascontiguousarray(rhs, np.float64)), _p64(x), C.byref(sp))
    return rc, x


def lu_sparse(topo: Topology, a_vals, rhs):
    x = np.zeros(topo.n)
    sp = C.c_double()
    rc = lib().orc_lu_sparse(topo.h, _p64(np.ascontiguousarray(a_vals, np.float64)),
                             _p64(np.ascontiguousarray(rhs, np.float64)), _p64(x), C.byref(sp))
    return rc, x


def lu_multifrontal(topo: Topology, a_vals, rhs):
    L = lib()
    L.orc_lu_multifrontal.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p]
    x = np.zeros(topo.n)
    sp = C.c_double()
    rc = L.orc_lu_multifrontal(topo.h, _p64(np.ascontiguousarray(a_vals, np.float64)),
                               _p64(np.ascontiguousarray(rhs, np.float64)), _p64(x), C.byref(sp))
    return rc, x


def lu_flag(topo: Topology, a_vals, rhs, x, bound=1e-8):
    """(flagged, eta): the AUTO route's test of a solve x, eta = max|b - A x| / max|b|."""
    L = lib()
    L.orc_lu_flag.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, C.c_double, C.POINTER(C.c_double)]
    w = C.c_double()
    f = L.orc_lu_flag(topo.h, _p64(np.ascontiguousarray(a_vals, np.float64)),
                      _p64(np.ascontiguousarray(rhs, np.float64)), _p64(np.ascontiguousarray(x, np.float64)),
                      bound, C.byref(w))
    return bool(f), w.value


def lu_auto(topo: Topology, a_vals, rhs):
    """(status, x, fallback): the multifrontal LU, redone by the dense
    pivoting LU when its residual is above 1e-8 of max|b| or it is singular."""
    L = lib()
    L.orc_lu_auto.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, C.POINTER(C.c_int32)]
    x = np.zeros(topo.n)
    fb = C.c_int32()
    rc = L.orc_lu_auto(topo.h, _p64(np.ascontiguousarray(a_vals, np.float64)),
                       _p64(np.ascontiguousarray(rhs, np.float64)), _p64(x), C.byref(fb))
    return rc, x, int(fb.value)


def lu_multifrontal_f32(topo: Topology, a_vals, rhs):
    L = lib()
    fp = C.POINTER(C.c_float)
    L.orc_lu_multifrontal_f32.argtypes = [C.c_void_p, fp, fp, fp]
    a = np.ascontiguousarray(a_vals, np.float32)
    r = np.ascontiguousarray(rhs, np.float32)
    x = np.zeros(topo.n, np.float32)
    rc = L.orc_lu_multifrontal_f32(topo.h, a.ctypes.data_as(fp), r.ctypes.data_as(fp), x.ctypes.data_as(fp))
    return rc, x


def jacobian_f32(nl, topo: Topology, v, gmin):
    """FP32 path: the float Jacobian (CSR values) of the DC system at v."""
    L = lib()
    fp = C.POINTER(C.c_float)
    L.orc_jacobian_f32.argtypes = [C.POINTER(_abi.NrNetlist), C.c_void_p, _f64p, C.c_double, fp]
    c, keep = nl.to_c()
    v = np.ascontiguousarray(v, np.float64)
    a = np.zeros(topo.nnz_a, np.float32)
    if L.orc_jacobian_f32(C.byref(c), topo.h, _p64(v), gmin, a.ctypes.data_as(fp)):
        raise RuntimeError("orc_jacobian_f32 failed")
    return a


def seed(nl):
    c, keep = nl.to_c()
    v = np.zeros(nl.num_nodes)
    lib().orc_seed(C.byref(c), _p64(v))
    return v


def pnjlim(vnew, vold, vt, vcrit):
    return lib().orc_pnjlim(vnew, vold, vt, vcrit)


def vcrit(vt, Is):
    return lib().orc_vcrit(vt, Is)
