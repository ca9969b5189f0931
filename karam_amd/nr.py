"""Python binding of the C ABI in include/karam_nr.h (libkaram_nr.so).

Mirrors the reference's operator surface for the NR hot path:
  - `Context`      GPU context (replaces the GPU context manager singleton)
  - `Topology`     one-time symbolic analysis (NetlistCompiler::compile +
                   symbolicFactorize, infrastructure/netlist_compiler.cpp:17-89)
  - `Batch`        B Monte-Carlo / corner instances of one topology, solved
                   together by CircuitSim::solveDC semantics
                   (solvers/circuitsim_solver.cpp:176-555)
  - `eval_diodes` / `eval_mosfets`   KernelDispatcher slot replacements
                   (tensors/kernel_dispatch.h:24-31)

The HIP library is the only compute path: if libkaram_nr.so is missing or no
GPU is visible, calls raise instead of falling back to anything on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkaram_nr.so")
_LIB = None

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p

# every symbol the C header declares (checked by the CPU test suite)
EXPORTED = [
    "nr_ctx_create", "nr_ctx_destroy", "nr_last_error", "nr_version",
    "nr_topology_build", "nr_topology_build_ex", "nr_topology_destroy", "nr_topology_get_info",
    "nr_topology_export",
    "nr_batch_alloc", "nr_batch_free", "nr_batch_set_params", "nr_batch_set_seed",
    "nr_batch_clear_seed", "nr_solve_opts_default", "nr_batch_solve_dc",
    "nr_batch_get_voltages", "nr_batch_get_stats", "nr_batch_get_timing", "nr_batch_get_kernel_stats",
    "nr_eval_diodes", "nr_eval_mosfets", "nr_eval_mosfets_noalloc",
    "nr_transient_opts_default", "nr_batch_transient", "nr_batch_get_waveform", "nr_batch_get_waveform_times",
    "nr_batch_get_wave_recorded", "nr_batch_get_tran_stats", "nr_batch_autotune",
    "nr_seeder_create", "nr_seeder_destroy", "nr_seeder_record", "nr_seeder_history", "nr_seeder_reset",
    "nr_train_opts_default", "nr_seeder_train", "nr_seeder_predict", "nr_batch_set_seed_from",
    "nr_seeder_param_count", "nr_seeder_get_params", "nr_seeder_set_params", "nr_seeder_loss_grad",
    "nr_spice_parse", "nr_spice_free", "nr_spice_get_netlist", "nr_spice_get_info", "nr_spice_node_name",
    "nr_spice_element_name", "nr_spice_warning", "nr_spice_vsrc_ac",
    "nr_ac_opts_default", "nr_batch_ac", "nr_batch_get_ac", "nr_batch_get_ac_freqs", "nr_batch_get_ac_status",
    "nr_batch_get_diagnostics", "nr_batch_get_trace", "nr_batch_set_launch_class", "nr_batch_set_persistent",
    "nr_batch_set_fanout",
    "nr_noise_opts_default", "nr_batch_noise", "nr_batch_get_noise", "nr_batch_get_noise_vrms",
    "nr_batch_get_noise_freqs", "nr_batch_get_noise_status",
]


class NrError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NrError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.nr_version.restype = C.c_char_p
        L.nr_last_error.restype = C.c_char_p
        L.nr_last_error.argtypes = [_vp]
        L.nr_ctx_create.argtypes = [C.POINTER(_abi.NrDeviceCfg), C.POINTER(_vp)]
        L.nr_ctx_destroy.argtypes = [_vp]
        L.nr_topology_build.argtypes = [_vp, C.POINTER(_abi.NrNetlist), C.POINTER(_vp)]
        L.nr_topology_build_ex.argtypes = [_vp, C.POINTER(_abi.NrNetlist), C.c_int32, C.POINTER(_vp)]
        L.nr_transient_opts_default.argtypes = [C.POINTER(_abi.NrTransientOpts)]
        L.nr_transient_opts_default.restype = None
        L.nr_batch_transient.argtypes = [_vp, C.POINTER(_abi.NrTransientOpts), C.POINTER(_abi.NrResult)]
        L.nr_batch_autotune.argtypes = [_vp, C.c_int32, C.c_int32, C.POINTER(_abi.NrAutotuneResult)]
        L.nr_batch_get_waveform.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_waveform_times.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_ac_opts_default.argtypes = [C.POINTER(_abi.NrAcOpts)]
        L.nr_ac_opts_default.restype = None
        L.nr_batch_ac.argtypes = [_vp, C.POINTER(_abi.NrAcOpts), C.POINTER(_abi.NrResult)]
        L.nr_batch_get_ac.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_ac_freqs.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_ac_status.argtypes = [_vp, C.POINTER(C.c_int32), C.c_int64]
        L.nr_batch_get_diagnostics.argtypes = [_vp, C.POINTER(_abi.NrDiagnostics), C.c_int64]
        L.nr_batch_set_launch_class.argtypes = [_vp] + [C.c_int32] * 5
        L.nr_batch_set_persistent.argtypes = [_vp, C.c_int32, C.c_int32]
        L.nr_batch_set_fanout.argtypes = [_vp, C.c_int32, C.c_int32]
        L.nr_batch_get_trace.argtypes = [_vp, C.POINTER(_abi.NrConvStep), C.c_int64]
        L.nr_noise_opts_default.argtypes = [C.POINTER(_abi.NrNoiseOpts)]
        L.nr_noise_opts_default.restype = None
        L.nr_batch_noise.argtypes = [_vp, C.POINTER(_abi.NrNoiseOpts), C.POINTER(_abi.NrResult)]
        L.nr_batch_get_noise.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_noise_vrms.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_noise_freqs.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_noise_status.argtypes = [_vp, C.POINTER(C.c_int32), C.c_int64]
        L.nr_spice_parse.argtypes = [C.c_char_p, C.c_int64, C.c_int32, C.POINTER(_vp)]
        L.nr_spice_free.argtypes = [_vp]
        L.nr_spice_get_netlist.argtypes = [_vp, C.POINTER(_abi.NrNetlist)]
        L.nr_spice_get_info.argtypes = [_vp, C.POINTER(_abi.NrSpiceInfo)]
        L.nr_spice_node_name.argtypes = [_vp, C.c_int32, C.c_char_p, C.c_int32]
        L.nr_spice_element_name.argtypes = [_vp, C.c_char, C.c_int32, C.c_char_p, C.c_int32]
        L.nr_spice_warning.argtypes = [_vp, C.c_int32, C.c_char_p, C.c_int32]
        L.nr_spice_vsrc_ac.argtypes = [_vp, C.c_int32, _f64p, _f64p]
        L.nr_batch_get_wave_recorded.argtypes = [_vp, C.POINTER(C.c_int64), C.c_int64]
        L.nr_batch_get_tran_stats.argtypes = [_vp, C.POINTER(_abi.NrTranStats), C.c_int64]
        _f32p = C.POINTER(C.c_float)
        L.nr_seeder_create.argtypes = [_vp, C.POINTER(_abi.NrSeederCfg), C.POINTER(_vp)]
        L.nr_seeder_destroy.argtypes = [_vp]
        L.nr_seeder_record.argtypes = [_vp, _f32p, _f64p, C.c_int64]
        L.nr_seeder_history.argtypes = [_vp]
        L.nr_seeder_history.restype = C.c_int64
        L.nr_seeder_reset.argtypes = [_vp]
        L.nr_train_opts_default.argtypes = [C.POINTER(_abi.NrTrainOpts)]
        L.nr_train_opts_default.restype = None
        L.nr_seeder_train.argtypes = [_vp, C.POINTER(_abi.NrTrainOpts), _f32p]
        L.nr_seeder_predict.argtypes = [_vp, _f32p, C.c_int64, _f64p]
        L.nr_batch_set_seed_from.argtypes = [_vp, _vp, _f32p]
        L.nr_seeder_param_count.argtypes = [_vp]
        L.nr_seeder_param_count.restype = C.c_int64
        L.nr_seeder_get_params.argtypes = [_vp, _f32p, C.c_int64]
        L.nr_seeder_set_params.argtypes = [_vp, _f32p, C.c_int64]
        L.nr_seeder_loss_grad.argtypes = [_vp, _f32p, _f32p, C.c_int64, _f32p, _f32p]
        L.nr_topology_destroy.argtypes = [_vp]
        L.nr_topology_get_info.argtypes = [_vp, C.POINTER(_abi.NrTopologyInfo)]
        L.nr_topology_export.argtypes = [_vp, C.c_int32, _i32p, C.c_int64, C.POINTER(C.c_int64)]
        L.nr_batch_alloc.argtypes = [_vp, _vp, C.c_int32, C.POINTER(_vp)]
        L.nr_batch_free.argtypes = [_vp]
        L.nr_batch_set_params.argtypes = [_vp, C.c_int32, _f64p, C.c_int64]
        L.nr_batch_set_seed.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_clear_seed.argtypes = [_vp]
        L.nr_solve_opts_default.argtypes = [C.POINTER(_abi.NrSolveOpts)]
        L.nr_solve_opts_default.restype = None
        L.nr_batch_solve_dc.argtypes = [_vp, C.POINTER(_abi.NrSolveOpts), C.POINTER(_abi.NrResult)]
        L.nr_batch_get_voltages.argtypes = [_vp, _f64p, C.c_int64]
        L.nr_batch_get_stats.argtypes = [_vp, C.POINTER(_abi.NrInstanceStats), C.c_int64]
        L.nr_batch_get_timing.argtypes = [_vp, C.POINTER(_abi.NrTiming)]
        L.nr_batch_get_kernel_stats.argtypes = [_vp, C.POINTER(_abi.NrKernelStat), C.c_int32,
                                                C.POINTER(C.c_int32)]
        L.nr_eval_diodes.argtypes = [_vp, C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, C.c_int32,
                                     _f64p, _f64p, _f64p]
        L.nr_eval_mosfets.argtypes = [_vp, C.c_int32, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                      _i32p, _f64p, C.c_int32, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.nr_eval_mosfets_noalloc.argtypes = L.nr_eval_mosfets.argtypes
        _LIB = L
    return _LIB


def _check(rc, ctx=None):
    if rc != _abi.NR_OK:
        msg = lib().nr_last_error(ctx.h if ctx is not None else None)
        raise NrError(f"karam_nr error {rc}: {msg.decode() if msg else ''}")


def _p64(a):
    return a.ctypes.data_as(_f64p)


def _p32(a):
    return a.ctypes.data_as(_i32p)


def _p32f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Seeder:
    """The learned seeder behind the NeuralSeeder seam
    (solvers/neural_seeder.h:45-220): record_solution / predict_seed /
    has_history / reset, plus train().  An MLP n_features -> n_hidden x
    hidden (GELU) -> n_out trained and evaluated on the GPU."""

    def __init__(self, ctx: Context, n_features: int, n_out: int, hidden: int = 1024, n_hidden: int = 3,
                 init_seed: int = 1):
        self.ctx, self.n_features, self.n_out = ctx, int(n_features), int(n_out)
        cfg = _abi.NrSeederCfg(self.n_features, self.n_out, int(hidden), int(n_hidden), int(init_seed))
        h = _vp()
        _check(lib().nr_seeder_create(ctx.h, C.byref(cfg), C.byref(h)), ctx)
        self.h = h

    def record_solution(self, features, v):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, self.n_features)
        vv = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, self.n_out)
        _check(lib().nr_seeder_record(self.h, _p32f(f), _p64(vv), f.shape[0]), self.ctx)

    def history(self) -> int:
        return int(lib().nr_seeder_history(self.h))

    def has_history(self) -> bool:
        return self.history() > 0

    def reset(self):
        _check(lib().nr_seeder_reset(self.h), self.ctx)

    def train(self, epochs=40, batch=256, lr=1e-3, shuffle_seed=1):
        o = _abi.NrTrainOpts()
        lib().nr_train_opts_default(C.byref(o))
        o.epochs, o.batch, o.lr, o.shuffle_seed = int(epochs), int(batch), float(lr), int(shuffle_seed)
        loss = C.c_float()
        _check(lib().nr_seeder_train(self.h, C.byref(o), C.byref(loss)), self.ctx)
        return float(loss.value)

    def predict_seed(self, features):
        f = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, self.n_features)
        out = np.zeros((f.shape[0], self.n_out))
        _check(lib().nr_seeder_predict(self.h, _p32f(f), f.shape[0], _p64(out)), self.ctx)
        return out

    def params(self):
        n = int(lib().nr_seeder_param_count(self.h))
        out = np.zeros(n, dtype=np.float32)
        _check(lib().nr_seeder_get_params(self.h, _p32f(out), n), self.ctx)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().nr_seeder_set_params(self.h, _p32f(p), p.size), self.ctx)

    def loss_grad(self, x, y):
        """Loss and flat gradient of one forward/backward pass (network units)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        n = int(lib().nr_seeder_param_count(self.h))
        g = np.zeros(n, dtype=np.float32)
        loss = C.c_float()
        _check(lib().nr_seeder_loss_grad(self.h, _p32f(x), _p32f(y), x.shape[0], _p32f(g), C.byref(loss)),
               self.ctx)
        return float(loss.value), g

    def close(self):
        if getattr(self, "h", None):
            lib().nr_seeder_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Context:
    def __init__(self, device: int = 0, verbose: int = 0):
        h = _vp()
        cfg = _abi.NrDeviceCfg(device, verbose)
        rc = lib().nr_ctx_create(C.byref(cfg), C.byref(h))
        _check(rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Topology:
    """Symbolic analysis of a netlist (host side; needs no GPU)."""

    def __init__(self, nl, ctx: Context | None = None, transient: bool = False):
        self.nl = nl
        self.transient = bool(transient)
        self._c, self._keep = nl.to_c()
        h = _vp()
        kind = _abi.NR_TOPO_TRANSIENT if transient else _abi.NR_TOPO_DC
        rc = lib().nr_topology_build_ex(ctx.h if ctx else None, C.byref(self._c), kind, C.byref(h))
        _check(rc, ctx)
        self.h = h
        info = _abi.NrTopologyInfo()
        _check(lib().nr_topology_get_info(h, C.byref(info)))
        self.info = {k: getattr(info, k) for k, _ in _abi.NrTopologyInfo._fields_}

    def export(self, what):
        n = C.c_int64()
        _check(lib().nr_topology_export(self.h, what, None, 0, C.byref(n)))
        buf = np.zeros(max(1, n.value), dtype=np.int32)
        _check(lib().nr_topology_export(self.h, what, _p32(buf), buf.size, C.byref(n)))
        return buf[: n.value].copy()

    def close(self):
        if getattr(self, "h", None):
            lib().nr_topology_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def solve_opts(max_iter=50, gmin_steps=19, precision=_abi.NR_PREC_FP64, seeded_start_step=18,
               record_timing=0, use_graph=0, fallback_plain=1, trace_capacity=0, lu=_abi.NR_LU_AUTO, polish=0):
    o = _abi.NrSolveOpts()
    lib().nr_solve_opts_default(C.byref(o))
    o.max_iter, o.gmin_steps, o.precision = max_iter, gmin_steps, precision
    o.seeded_start_step, o.record_timing, o.use_graph = seeded_start_step, record_timing, use_graph
    o.fallback_plain, o.trace_capacity, o.lu, o.polish = fallback_plain, trace_capacity, int(lu), int(polish)
    return o


class Batch:
    """B instances of one topology; parameters per instance (MC / corners)."""

    def __init__(self, ctx: Context, topo: Topology, B: int):
        self.ctx, self.topo, self.B = ctx, topo, int(B)
        h = _vp()
        _check(lib().nr_batch_alloc(ctx.h, topo.h, self.B, C.byref(h)), ctx)
        self.h = h

    def set_params(self, fid: int, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().nr_batch_set_params(self.h, fid, _p64(a), a.size), self.ctx)

    def set_fields(self, fields: dict):
        for fid, arr in fields.items():
            self.set_params(fid, np.asarray(arr)[: self.B])

    def set_seed(self, v0):
        a = np.ascontiguousarray(v0, dtype=np.float64)
        _check(lib().nr_batch_set_seed(self.h, _p64(a), a.size), self.ctx)

    def set_seed_from(self, seeder: "Seeder", features):
        """The seed of every instance predicted by the seeder on the device."""
        f = np.ascontiguousarray(features, dtype=np.float32)
        _check(lib().nr_batch_set_seed_from(self.h, seeder.h, _p32f(f)), self.ctx)

    def clear_seed(self):
        _check(lib().nr_batch_clear_seed(self.h), self.ctx)

    def solve_dc(self, opts=None, **kw):
        o = opts if opts is not None else solve_opts(**kw)
        r = _abi.NrResult()
        _check(lib().nr_batch_solve_dc(self.h, C.byref(o), C.byref(r)), self.ctx)
        self._trace_cap = int(o.trace_capacity)
        return {k: getattr(r, k) for k, _ in _abi.NrResult._fields_}

    def diagnostics(self):
        """SolverStats diagnostics of the last DC solve, per instance:
        {history, failure (NR_FAIL_*), stagnation, divergence, worst_node,
        worst_node_v, rms_residual, gmin_stepping_used} arrays."""
        arr = (_abi.NrDiagnostics * self.B)()
        _check(lib().nr_batch_get_diagnostics(self.h, arr, self.B), self.ctx)
        return {k: np.array([getattr(x, k) for x in arr]) for k, _ in _abi.NrDiagnostics._fields_ if k != "pad"}

    def trace(self):
        """The last solve's ConvergenceStep rings: {field: [B, capacity]}
        (record j of an instance at j mod capacity)."""
        cap = getattr(self, "_trace_cap", 0)
        arr = (_abi.NrConvStep * (self.B * cap))()
        _check(lib().nr_batch_get_trace(self.h, arr, self.B * cap), self.ctx)
        return {k: np.array([getattr(x, k) for x in arr]).reshape(self.B, cap)
                for k, _ in _abi.NrConvStep._fields_ if k != "pad"}

    def transient(self, h=1e-9, steps=200, max_iter=50, start=_abi.NR_TRAN_START_ZERO, probes=None,
                  record_timing=0, precision=_abi.NR_PREC_FP64, use_graph=0, method=_abi.NR_INT_GEAR1,
                  control=_abi.NR_TRAN_FIXED, t_stop=0.0, wave_capacity=0, wave_decimation=1, lu=_abi.NR_LU_AUTO):
        """Transient of every instance (needs a transient topology): `steps`
        fixed steps of h, or with control=NR_TRAN_ARBITER the reference's
        step control up to t_stop.  probes: 1-based node ids recorded in the
        waveform ring."""
        o = _abi.NrTransientOpts()
        lib().nr_transient_opts_default(C.byref(o))
        self._probes = np.ascontiguousarray(probes if probes is not None else [], dtype=np.int32)
        o.h, o.steps, o.max_iter, o.start = float(h), int(steps), int(max_iter), int(start)
        o.n_probe = len(self._probes)
        o.probe = _p32(self._probes) if len(self._probes) else None
        o.record_timing = record_timing
        o.precision = int(precision)
        o.use_graph = int(use_graph)
        o.method, o.control, o.t_stop = int(method), int(control), float(t_stop)
        o.wave_capacity, o.wave_decimation = int(wave_capacity), int(wave_decimation)
        o.lu = int(lu)
        r = _abi.NrResult()
        _check(lib().nr_batch_transient(self.h, C.byref(o), C.byref(r)), self.ctx)
        ring = int(wave_capacity)
        if ring == 0:
            ring = (int(np.ceil(float(t_stop) / float(h))) + 1 if int(control) == _abi.NR_TRAN_ARBITER
                    else int(steps) + 1)
        self._wave_shape = (self.B, ring, len(self._probes))
        return {k: getattr(r, k) for k, _ in _abi.NrResult._fields_}

    def ac(self, f_start=1.0, f_stop=1e9, n_points=50, input_node=0, amplitude=1.0, probes=None, record_timing=0):
        """AC sweep of every instance at its operating point (the seed);
        returns (freqs [F], H [B, F, n_probe] complex, status [B])."""
        o = _abi.NrAcOpts()
        lib().nr_ac_opts_default(C.byref(o))
        pr = np.ascontiguousarray(probes if probes is not None else [], dtype=np.int32)
        o.f_start, o.f_stop, o.n_points = float(f_start), float(f_stop), int(n_points)
        o.input_node, o.amplitude = int(input_node), float(amplitude)
        o.n_probe = len(pr)
        o.probe = _p32(pr) if len(pr) else None
        o.record_timing = record_timing
        r = _abi.NrResult()
        _check(lib().nr_batch_ac(self.h, C.byref(o), C.byref(r)), self.ctx)
        npr = len(pr) if len(pr) else self.topo.info["n"]
        out = np.zeros((self.B, int(n_points), npr, 2))
        _check(lib().nr_batch_get_ac(self.h, _p64(out), out.size), self.ctx)
        freqs = np.zeros(int(n_points))
        _check(lib().nr_batch_get_ac_freqs(self.h, _p64(freqs), freqs.size), self.ctx)
        st = np.zeros(self.B, dtype=np.int32)
        _check(lib().nr_batch_get_ac_status(self.h, st.ctypes.data_as(C.POINTER(C.c_int32)), self.B), self.ctx)
        self._ac_result = {k: getattr(r, k) for k, _ in _abi.NrResult._fields_}
        return freqs, out[..., 0] + 1j * out[..., 1], st

    def noise(self, output_node, f_start=1.0, f_stop=1e9, n_points=50, temp_K=300.15, input_node=0,
              record_timing=0):
        """Noise analysis of every instance at its operating point (the seed);
        returns (freqs [F], output noise PSD [B, F] in V^2/Hz, integrated RMS
        noise [B] in V, status [B])."""
        o = _abi.NrNoiseOpts()
        lib().nr_noise_opts_default(C.byref(o))
        o.f_start, o.f_stop, o.n_points = float(f_start), float(f_stop), int(n_points)
        o.output_node, o.temp_K, o.input_node = int(output_node), float(temp_K), int(input_node)
        o.record_timing = record_timing
        r = _abi.NrResult()
        _check(lib().nr_batch_noise(self.h, C.byref(o), C.byref(r)), self.ctx)
        psd = np.zeros((self.B, int(n_points)))
        _check(lib().nr_batch_get_noise(self.h, _p64(psd), psd.size), self.ctx)
        vrms = np.zeros(self.B)
        _check(lib().nr_batch_get_noise_vrms(self.h, _p64(vrms), self.B), self.ctx)
        freqs = np.zeros(int(n_points))
        _check(lib().nr_batch_get_noise_freqs(self.h, _p64(freqs), freqs.size), self.ctx)
        st = np.zeros(self.B, dtype=np.int32)
        _check(lib().nr_batch_get_noise_status(self.h, st.ctypes.data_as(C.POINTER(C.c_int32)), self.B), self.ctx)
        self._noise_result = {k: getattr(r, k) for k, _ in _abi.NrResult._fields_}
        return freqs, psd, vrms, st

    def set_launch_class(self, bucket, team, block, full=0, precision=_abi.NR_PREC_FP64):
        """Launch class of one front-order bucket (nr_batch_set_launch_class)."""
        _check(lib().nr_batch_set_launch_class(self.h, int(precision), int(bucket), int(team), int(block),
                                               int(full)), self.ctx)

    def set_persistent(self, on=True, precision=_abi.NR_PREC_FP64):
        """The large-front phase in one launch each for the factorisation and
        the back substitution (nr_batch_set_persistent); kept by the topology."""
        _check(lib().nr_batch_set_persistent(self.h, int(precision), int(bool(on))), self.ctx)

    def set_fanout(self, on=True, precision=_abi.NR_PREC_FP64):
        """The launches of one elimination-tree level on concurrent streams
        (nr_batch_set_fanout); kept by the topology."""
        _check(lib().nr_batch_set_fanout(self.h, int(precision), int(bool(on))), self.ctx)

    def autotune(self, precision=_abi.NR_PREC_FP64, reps=3):
        """Choose the large-front launch shape per front-order bucket by timing
        the candidates on this batch (nr_batch_autotune); the choice is kept
        by the topology.  Returns the chosen map and the factor + back
        substitution time before and after."""
        r = _abi.NrAutotuneResult()
        _check(lib().nr_batch_autotune(self.h, int(precision), int(reps), C.byref(r)), self.ctx)
        return {"buckets": [{"rmax": r.bucket_rmax[b], "team": r.team[b], "block": r.block[b], "fronts": r.fronts[b],
                             "full": r.full[b]} for b in range(r.n_buckets)],
                "ms_before": r.ms_before, "ms_after": r.ms_after, "ms_levels": r.ms_levels,
                "ms_persistent": r.ms_persistent, "persistent": r.persistent, "ms_graph": r.ms_graph,
                "ms_graph_fanout": r.ms_graph_fanout, "fanout": r.fanout, "ms_few_before": r.ms_few_before,
                "ms_few_after": r.ms_few_after, "few_max": r.few_max}

    def waveform(self):
        """Probe voltages of the last transient's waveform ring,
        [B, capacity, n_probe] (entry k in slot k mod capacity)."""
        out = np.zeros(self._wave_shape)
        _check(lib().nr_batch_get_waveform(self.h, _p64(out), out.size), self.ctx)
        return out

    def waveform_times(self):
        """Times of the waveform ring's entries, [B, capacity]."""
        out = np.zeros(self._wave_shape[:2])
        _check(lib().nr_batch_get_waveform_times(self.h, _p64(out), out.size), self.ctx)
        return out

    def wave_recorded(self):
        """Entries written to each instance's waveform ring."""
        out = np.zeros(self.B, dtype=np.int64)
        _check(lib().nr_batch_get_wave_recorded(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)), self.B),
               self.ctx)
        return out

    def tran_stats(self):
        """Per-instance step control of the last transient."""
        arr = (_abi.NrTranStats * self.B)()
        _check(lib().nr_batch_get_tran_stats(self.h, arr, self.B), self.ctx)
        return {k: np.array([getattr(x, k) for x in arr]) for k, _ in _abi.NrTranStats._fields_ if k != "pad"}

    def voltages(self):
        n = self.topo.info["n"]
        out = np.zeros((self.B, n))
        _check(lib().nr_batch_get_voltages(self.h, _p64(out), out.size), self.ctx)
        return out

    def stats(self):
        arr = (_abi.NrInstanceStats * self.B)()
        _check(lib().nr_batch_get_stats(self.h, arr, self.B), self.ctx)
        return {
            "iterations": np.array([x.iterations for x in arr]),
            "ref_iterations": np.array([x.ref_iterations for x in arr]),
            "status": np.array([x.status for x in arr]),
            "gmin_steps_done": np.array([x.gmin_steps_done for x in arr]),
            "kcl": np.array([x.kcl for x in arr]),
            "seeded": np.array([x.seeded for x in arr]),
            "lu_fallbacks": np.array([x.lu_fallbacks for x in arr]),
        }

    def timing(self):
        t = _abi.NrTiming()
        _check(lib().nr_batch_get_timing(self.h, C.byref(t)), self.ctx)
        return {"ms": list(t.ms), "launches": list(t.launches), "factor_kernel_ms": t.factor_kernel_ms,
                "factor_instances": t.factor_instances}

    def kernel_stats(self):
        """Per-kernel rows of the last solve run with record_timing=2:
        {name: {launches, ms, bytes, flops}} (bytes and flops algorithmic)."""
        n = C.c_int32()
        _check(lib().nr_batch_get_kernel_stats(self.h, None, 0, C.byref(n)), self.ctx)
        arr = (_abi.NrKernelStat * max(1, n.value))()
        _check(lib().nr_batch_get_kernel_stats(self.h, arr, n.value, C.byref(n)), self.ctx)
        out = {}
        for r in arr[: n.value]:       # two families can share a kernel (the back substitutions): add them up
            x = out.setdefault(r.name.decode(), {"launches": 0, "ms": 0.0, "bytes": 0.0, "flops": 0.0})
            x["launches"] += r.launches
            x["ms"] += r.ms
            x["bytes"] += r.bytes
            x["flops"] += r.flops
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().nr_batch_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def eval_diodes(ctx: Context, a, c, Is, N, Vt, v):
    """KernelDispatcher::batch_diode_physics replacement (host arrays)."""
    a, c = np.ascontiguousarray(a, np.int32), np.ascontiguousarray(c, np.int32)
    Is, N, Vt, v = (np.ascontiguousarray(x, np.float64) for x in (Is, N, Vt, v))
    m = len(a)
    vd, i_d, gd = np.zeros(m), np.zeros(m), np.zeros(m)
    _check(lib().nr_eval_diodes(ctx.h, m, _p32(a), _p32(c), _p64(Is), _p64(N), _p64(Vt), _p64(v), len(v),
                                _p64(vd), _p64(i_d), _p64(gd)), ctx)
    return vd, i_d, gd


def eval_mosfets(ctx: Context, d, g, s, W, L, Kp, Vth, lam, pmos, v):
    """KernelDispatcher::batch_mosfet_physics replacement (host arrays)."""
    d, g, s, pmos = (np.ascontiguousarray(x, np.int32) for x in (d, g, s, pmos))
    W, L, Kp, Vth, lam, v = (np.ascontiguousarray(x, np.float64) for x in (W, L, Kp, Vth, lam, v))
    m = len(d)
    outs = [np.zeros(m) for _ in range(5)]
    _check(lib().nr_eval_mosfets(ctx.h, m, _p32(d), _p32(g), _p32(s), _p64(W), _p64(L), _p64(Kp), _p64(Vth),
                                 _p64(lam), _p32(pmos), _p64(v), len(v), *[_p64(o) for o in outs]), ctx)
    return tuple(outs)
