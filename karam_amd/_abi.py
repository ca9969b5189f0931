"""ctypes mirror of include/karam_nr.h (structs, enums, field ids).

The netlist structure is the wire format between callers and the C ABI; it
follows the reference's component structs (netlist/circuit.h:31-41, 143-171,
392-408, 467-505).  Nothing here touches a GPU.
"""
import ctypes as C

NR_OK = 0
NR_ERR_INVALID = 1
NR_ERR_HIP = 2
NR_ERR_NO_DEVICE = 3
NR_ERR_TOPOLOGY = 4
NR_ERR_OOM = 5
NR_ERR_UNSUPPORTED = 6

NR_FIELD_RES_OHMS = 0
NR_FIELD_DIO_IS = 1
NR_FIELD_MOS_KP = 2
NR_FIELD_MOS_VTO = 3
NR_FIELD_ISRC_AMPS = 4
NR_FIELD_VSRC_VOLTS = 5
NR_FIELD_CAP_FARADS = 6
NR_FIELD_TEMP_K = 7
N_FIELDS = 8

NR_EXPORT_ROW_PTR = 0
NR_EXPORT_COL_IDX = 1
NR_EXPORT_STAMP_SLOT = 2
NR_EXPORT_ORDERING_RAW = 3
NR_EXPORT_PERM = 4
NR_EXPORT_FILL_ROW_PTR = 5
NR_EXPORT_FILL_COL_IDX = 6
NR_EXPORT_DIODE_ORDER = 7
NR_EXPORT_MOS_ORDER = 8
NR_EXPORT_SN_FIRST = 9

NR_PREC_FP64 = 0
NR_PREC_FP32_REFINE = 1
# linear-solve routes (nr_solve_opts.lu / nr_transient_opts.lu)
NR_LU_AUTO, NR_LU_MULTIFRONTAL, NR_LU_DENSE = 0, 1, 2
NR_DENSE_MAX_N = 4096

NR_STATUS_CONVERGED = 0
NR_STATUS_NOT_CONVERGED = 1
NR_STATUS_SINGULAR = 2

NR_PHASE_EVAL_STAMP = 0
NR_PHASE_FACTOR = 1
NR_PHASE_BACKSOLVE = 2
NR_PHASE_UPDATE = 3
NR_PHASE_COUNT = 4

NR_TOPO_DC = 0
NR_TOPO_TRANSIENT = 1
NR_TRAN_START_ZERO = 0
NR_TRAN_START_SEED = 1
NR_WAVE_DC = 0
NR_WAVE_PULSE = 1
NR_WAVE_SINE = 2
NR_INT_TRAP = 0
NR_INT_GEAR1 = 1
NR_INT_GEAR2 = 2
NR_TRAN_FIXED = 0
NR_TRAN_ARBITER = 1
NR_SPICE_VSS_IS_NET = 1

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class NrModelCard(C.Structure):
    _fields_ = [
        ("name", C.c_char_p), ("type", C.c_char_p),
        ("has_kp", C.c_int32), ("has_vto", C.c_int32), ("has_lambda", C.c_int32),
        ("kp", C.c_double), ("vto", C.c_double), ("lam", C.c_double),
    ]


class NrNetlist(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int32), ("ambient_temp_K", C.c_double),
        ("n_res", C.c_int32), ("res_n1", _i32p), ("res_n2", _i32p), ("res_ohms", _f64p),
        ("n_cap", C.c_int32), ("cap_n1", _i32p), ("cap_n2", _i32p), ("cap_farads", _f64p),
        ("n_vsrc", C.c_int32), ("vsrc_pos", _i32p), ("vsrc_neg", _i32p), ("vsrc_volts", _f64p),
        ("vsrc_wave", _i32p), ("vsrc_wave_par", _f64p),
        ("n_isrc", C.c_int32), ("isrc_pos", _i32p), ("isrc_neg", _i32p), ("isrc_amps", _f64p),
        ("isrc_wave", _i32p), ("isrc_wave_par", _f64p),
        ("n_diode", C.c_int32), ("dio_anode", _i32p), ("dio_cathode", _i32p),
        ("dio_is", _f64p), ("dio_n", _f64p), ("dio_vt", _f64p), ("dio_model", _i32p),
        ("n_mos", C.c_int32), ("mos_d", _i32p), ("mos_g", _i32p), ("mos_s", _i32p), ("mos_b", _i32p),
        ("mos_w", _f64p), ("mos_l", _f64p), ("mos_model", _i32p), ("mos_kp", _f64p), ("mos_vto", _f64p),
        ("n_models", C.c_int32), ("models", C.POINTER(NrModelCard)),
    ]


class NrDeviceCfg(C.Structure):
    _fields_ = [("device", C.c_int32), ("verbose", C.c_int32)]


class NrTopologyInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nnz_a", C.c_int32), ("nnz_f", C.c_int32), ("n_stamps", C.c_int32),
        ("treewidth", C.c_int32), ("ordering_len", C.c_int32), ("fill_in", C.c_int32),
        ("n_supernodes", C.c_int32), ("n_levels", C.c_int32), ("max_front", C.c_int32),
        ("n_res", C.c_int32), ("n_vsrc", C.c_int32), ("n_isrc", C.c_int32), ("n_cap", C.c_int32),
        ("n_diode", C.c_int32), ("n_mos", C.c_int32),
        ("lu_flops", C.c_int64), ("u_entries", C.c_int64), ("analyse_ms", C.c_double),
    ]


class NrSolveOpts(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32), ("gmin_steps", C.c_int32), ("precision", C.c_int32),
        ("seeded_start_step", C.c_int32), ("record_timing", C.c_int32), ("use_graph", C.c_int32),
        ("fallback_plain", C.c_int32), ("trace_capacity", C.c_int32), ("lu", C.c_int32), ("polish", C.c_int32),
    ]


class NrDiagnostics(C.Structure):
    _fields_ = [
        ("history", C.c_int64), ("failure", C.c_int32), ("stagnation", C.c_int32), ("divergence", C.c_int32),
        ("worst_node", C.c_int32), ("worst_node_v", C.c_double), ("rms_residual", C.c_double),
        ("gmin_stepping_used", C.c_int32), ("pad", C.c_int32),
    ]


class NrConvStep(C.Structure):
    _fields_ = [
        ("residual", C.c_double), ("max_dv", C.c_double), ("damping", C.c_double),
        ("gmin_step", C.c_int32), ("iteration", C.c_int32), ("worst_node", C.c_int32), ("pad", C.c_int32),
    ]


NR_FAIL_NONE, NR_FAIL_SINGULAR, NR_FAIL_DIVERGENCE, NR_FAIL_LIMIT_CYCLE, NR_FAIL_SLOW, NR_FAIL_FLOATING_NODE, \
    NR_FAIL_OVERFLOW = range(7)


class NrTransientOpts(C.Structure):
    _fields_ = [
        ("h", C.c_double), ("steps", C.c_int32), ("max_iter", C.c_int32), ("start", C.c_int32),
        ("n_probe", C.c_int32), ("probe", C.POINTER(C.c_int32)), ("record_timing", C.c_int32),
        ("precision", C.c_int32), ("use_graph", C.c_int32), ("method", C.c_int32), ("control", C.c_int32),
        ("t_stop", C.c_double), ("wave_capacity", C.c_int32), ("wave_decimation", C.c_int32), ("lu", C.c_int32),
    ]


class NrTranStats(C.Structure):
    _fields_ = [
        ("t_end", C.c_double), ("h_last", C.c_double), ("accepted", C.c_int32), ("rejected", C.c_int32),
        ("method", C.c_int32), ("switched_at", C.c_int32), ("instability", C.c_int32), ("pad", C.c_int32),
    ]


class NrSeederCfg(C.Structure):
    _fields_ = [("n_features", C.c_int32), ("n_out", C.c_int32), ("hidden", C.c_int32), ("n_hidden", C.c_int32),
                ("init_seed", C.c_uint64)]


class NrTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("shuffle_seed", C.c_uint64)]


class NrResult(C.Structure):
    _fields_ = [
        ("B", C.c_int32), ("n_converged", C.c_int32), ("batched_iterations", C.c_int32),
        ("max_iterations", C.c_int32), ("mean_iterations", C.c_double), ("max_kcl", C.c_double),
        ("solve_ms", C.c_double), ("device_evals", C.c_double),
    ]


class NrInstanceStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("ref_iterations", C.c_int32), ("status", C.c_int32),
        ("gmin_steps_done", C.c_int32), ("kcl", C.c_double),
        ("seeded", C.c_int32), ("lu_fallbacks", C.c_int32),
    ]


NR_KERNEL_NAME_LEN = 48


class NrKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * NR_KERNEL_NAME_LEN), ("launches", C.c_int64), ("ms", C.c_double),
                ("bytes", C.c_double), ("flops", C.c_double)]


NR_TUNE_BUCKETS = 8


class NrAutotuneResult(C.Structure):
    _fields_ = [("n_buckets", C.c_int32), ("bucket_rmax", C.c_int32 * NR_TUNE_BUCKETS),
                ("team", C.c_int32 * NR_TUNE_BUCKETS), ("block", C.c_int32 * NR_TUNE_BUCKETS),
                ("fronts", C.c_int32 * NR_TUNE_BUCKETS), ("full", C.c_int32 * NR_TUNE_BUCKETS),
                ("ms_before", C.c_double), ("ms_after", C.c_double), ("ms_levels", C.c_double),
                ("ms_persistent", C.c_double), ("persistent", C.c_int32), ("ms_graph", C.c_double),
                ("ms_graph_fanout", C.c_double), ("fanout", C.c_int32), ("ms_few_before", C.c_double),
                ("ms_few_after", C.c_double), ("few_max", C.c_int32)]


class NrTiming(C.Structure):
    _fields_ = [
        ("ms", C.c_double * NR_PHASE_COUNT), ("launches", C.c_int64 * NR_PHASE_COUNT),
        ("factor_kernel_ms", C.c_double), ("factor_instances", C.c_int64),
    ]


class NrSpiceInfo(C.Structure):
    _fields_ = [
        ("lines", C.c_int32), ("elements", C.c_int32), ("models", C.c_int32), ("nodes", C.c_int32),
        ("has_op", C.c_int32), ("has_tran", C.c_int32), ("has_ac", C.c_int32), ("has_dc", C.c_int32),
        ("tran_step", C.c_double), ("tran_stop", C.c_double), ("tran_start", C.c_double),
        ("warnings", C.c_int32), ("pad", C.c_int32),
    ]


class NrNoiseOpts(C.Structure):
    _fields_ = [
        ("f_start", C.c_double), ("f_stop", C.c_double), ("n_points", C.c_int32), ("output_node", C.c_int32),
        ("temp_K", C.c_double), ("input_node", C.c_int32), ("record_timing", C.c_int32),
    ]


class NrAcOpts(C.Structure):
    _fields_ = [
        ("f_start", C.c_double), ("f_stop", C.c_double), ("n_points", C.c_int32), ("input_node", C.c_int32),
        ("amplitude", C.c_double), ("n_probe", C.c_int32), ("probe", C.POINTER(C.c_int32)),
        ("record_timing", C.c_int32), ("pad", C.c_int32),
    ]
