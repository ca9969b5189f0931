/*
 * karam_nr.h — C ABI of the MI355X Newton-Raphson inner loop.
 *
 * Drop-in boundary for the reference engine's NR hot path.  Every entry point
 * is `extern "C"`, takes plain pointers and sizes, returns an int status
 * (NR_OK == 0) and never lets an exception cross the boundary.  The exported
 * symbols carry default visibility (the reference's ENGINE_API convention,
 * engine_export.h:19-39); everything else in the library is hidden.
 *
 * Conventions (mirroring the reference):
 *   - node ids are 1-based, 0 is ground; `num_nodes` is numGlobalNodes
 *     (netlist/circuit.h:1056-1267);
 *   - voltages vectors are 0-based, index k holds node k+1;
 *   - component arrays are given in netlist order (the order the caller
 *     added them); the library derives the reference's SoA device order
 *     itself (tensorizeNetlist, tensors/physics_tensors.cpp:305-493);
 *   - caller-owned host buffers, library-owned device buffers, opaque
 *     handles; a batch is not re-entrant, different batches may be driven
 *     from different host threads.
 */
#ifndef KARAM_NR_H
#define KARAM_NR_H

#include <stdint.h>

#if defined(KARAM_NR_BUILD)
#define NR_API __attribute__((visibility("default")))
#else
#define NR_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirror SolverResult / SolverStats semantics) -------- */
enum {
    NR_OK = 0,
    NR_ERR_INVALID = 1,      /* bad argument / shape                         */
    NR_ERR_HIP = 2,          /* HIP runtime failure (message in nr_last_error)*/
    NR_ERR_NO_DEVICE = 3,    /* no GPU visible                               */
    NR_ERR_TOPOLOGY = 4,     /* netlist cannot be compiled (e.g. node gaps)  */
    NR_ERR_OOM = 5,          /* device allocation failed                     */
    NR_ERR_UNSUPPORTED = 6   /* feature not provided by this build           */
};

/* ---- netlist (caller-owned, netlist order) ------------------------------- */

/* Transient waveforms of independent sources (VoltageSource / CurrentSource
 * type, netlist/circuit.h:143-171; evaluated as stampSoABlock does,
 * physics/circuitsim.h:2150-2240).  A DC solve uses the DC value.
 *   NR_WAVE_PULSE: v1 v2 td tr tf pw per
 *   NR_WAVE_SINE:  vo va freq td theta phase_deg (7th value unused)      */
enum { NR_WAVE_DC = 0, NR_WAVE_PULSE = 1, NR_WAVE_SINE = 2 };

/* A `.model` card (netlist/circuit.h:31-41).  `type` is "NMOS", "PMOS" or
 * "D"; has_* flags say whether the card sets the parameter (ModelCard::get
 * falls back to the tensorizer defaults otherwise). */
typedef struct nr_model_card {
    const char* name;
    const char* type;
    int32_t has_kp, has_vto, has_lambda;
    double kp, vto, lambda;
} nr_model_card;

typedef struct nr_netlist {
    int32_t num_nodes;          /* numGlobalNodes                            */
    double ambient_temp_K;      /* SimulationEnvironment, default 300.15     */

    int32_t n_res;              /* Resistor (circuit.h)                      */
    const int32_t* res_n1;
    const int32_t* res_n2;
    const double* res_ohms;

    int32_t n_cap;              /* Capacitor                                 */
    const int32_t* cap_n1;
    const int32_t* cap_n2;
    const double* cap_farads;

    int32_t n_vsrc;             /* VoltageSource (circuit.h:143-171)         */
    const int32_t* vsrc_pos;
    const int32_t* vsrc_neg;
    const double* vsrc_volts;   /* DC value (voltage_V)                      */
    const int32_t* vsrc_wave;   /* nullable: NR_WAVE_* per source            */
    const double* vsrc_wave_par; /* nullable: 7 per source, see NR_WAVE_*    */

    int32_t n_isrc;             /* CurrentSource                             */
    const int32_t* isrc_pos;
    const int32_t* isrc_neg;
    const double* isrc_amps;    /* DC value (current_A)                      */
    const int32_t* isrc_wave;   /* nullable: NR_WAVE_* per source            */
    const double* isrc_wave_par; /* nullable: 7 per source, see NR_WAVE_*    */

    int32_t n_diode;            /* Diode (circuit.h:392-408)                 */
    const int32_t* dio_anode;
    const int32_t* dio_cathode;
    const double* dio_is;
    const double* dio_n;
    const double* dio_vt;       /* thermalVoltage_V_T_V (unscaled)           */
    const int32_t* dio_model;   /* nullable: index into `models` for the model
                                   name; -1 / NULL = "diode_generic"          */

    int32_t n_mos;              /* Mosfet (circuit.h:467-505)                */
    const int32_t* mos_d;
    const int32_t* mos_g;
    const int32_t* mos_s;
    const int32_t* mos_b;
    const double* mos_w;
    const double* mos_l;
    const int32_t* mos_model;   /* index into `models` (model name + card)   */
    const double* mos_kp;       /* nullable per-device KP override (MC)      */
    const double* mos_vto;      /* nullable per-device VTO override (MC)     */

    int32_t n_models;
    const nr_model_card* models; /* model cards; an entry whose `type` is NULL
                                    names a model with no card               */
} nr_netlist;

/* ---- SPICE netlist ingestion -------------------------------------------- */
/* SPICEParser::parse (netlist/spice_parser.cpp:148-261, the entry the
 * session's deserialiseRequest calls, engine_impl.cpp:84-95): SPICE text to
 * an nr_netlist with the reference parser's node numbering (first
 * appearance; 0, GND, GROUND and VSS are ground) and component order.
 * Filled in where that parser leaves the netlist undefined or drops what the
 * line says: diode Is / N / Vt from the diode's .model card (defaults 1e-14
 * A, 1, 0.02585 V); a PULSE / SIN source without a DC value takes its t = 0
 * value; SIN / SINE select NR_WAVE_SINE; MOSFET W / L are read; AC on a
 * current source is skipped.  Elements the batched Newton path does not
 * simulate (L, K, Q, J, E, G, H, F, T, .subckt / X) make the parse fail with
 * NR_ERR_UNSUPPORTED and the line number in nr_last_error(NULL).  Unknown
 * element letters are skipped with a warning, as the reference does. */
typedef struct nr_spice nr_spice;
typedef struct nr_spice_info {
    int32_t lines;           /* lines after the title (SPICEParseResult)    */
    int32_t elements;        /* element lines, skipped ones included        */
    int32_t models;          /* .model cards                                */
    int32_t nodes;           /* numGlobalNodes                              */
    int32_t has_op, has_tran, has_ac, has_dc;
    double tran_step, tran_stop, tran_start;
    int32_t warnings;
    int32_t pad;
} nr_spice_info;
enum { NR_SPICE_VSS_IS_NET = 1 };   /* flag: VSS is an ordinary net         */
NR_API int nr_spice_parse(const char* text, int64_t len, int32_t flags, nr_spice** out);
NR_API int nr_spice_free(nr_spice* p);
/* A view of the parsed netlist: its arrays stay valid until nr_spice_free. */
NR_API int nr_spice_get_netlist(const nr_spice* p, nr_netlist* out);
NR_API int nr_spice_get_info(const nr_spice* p, nr_spice_info* out);
/* Names, NUL-terminated into buf (cap bytes): node k (upper-cased, as the
 * parser keys it); element `index` of kind 'R' 'C' 'V' 'I' 'D' 'M' (its
 * first token); warning k. */
NR_API int nr_spice_node_name(const nr_spice* p, int32_t node, char* buf, int32_t cap);
NR_API int nr_spice_element_name(const nr_spice* p, char kind, int32_t index, char* buf, int32_t cap);
NR_API int nr_spice_warning(const nr_spice* p, int32_t k, char* buf, int32_t cap);
/* AC magnitude / phase of voltage source i (read, not simulated here). */
NR_API int nr_spice_vsrc_ac(const nr_spice* p, int32_t i, double* mag, double* phase_deg);

/* ---- per-instance parameter fields (Monte-Carlo / corners) -------------- */
/* nr_batch_set_params(field, values[B * count]) — row-major per instance,
 * netlist order, absolute values (already perturbed). */
enum {
    NR_FIELD_RES_OHMS = 0,   /* [B][n_res]   */
    NR_FIELD_DIO_IS = 1,     /* [B][n_diode] */
    NR_FIELD_MOS_KP = 2,     /* [B][n_mos]   */
    NR_FIELD_MOS_VTO = 3,    /* [B][n_mos]   */
    NR_FIELD_ISRC_AMPS = 4,  /* [B][n_isrc]  */
    NR_FIELD_VSRC_VOLTS = 5, /* [B][n_vsrc]  */
    NR_FIELD_CAP_FARADS = 6, /* [B][n_cap]   */
    /* [B][1]: each instance's ambient temperature (K), a corner's
     * temperature (CornerConfig::temperatureC, physics/corner_config.h:42-95).
     * It scales the diodes' V_T as tensorizeNetlist does
     * (tensors/physics_tensors.cpp:419) and again in pnjlim
     * (solvers/circuitsim_solver.cpp:127-129); unset, every instance takes
     * the netlist's ambient_temp_K.  Level-1 MOSFETs have no temperature
     * terms in the reference. */
    NR_FIELD_TEMP_K = 7
};

/* ---- opaque handles ------------------------------------------------------ */
typedef struct nr_ctx nr_ctx;
typedef struct nr_topology nr_topology;
typedef struct nr_batch nr_batch;

typedef struct nr_device_cfg {
    int32_t device;          /* HIP device ordinal                           */
    int32_t verbose;         /* 0 = silent                                   */
} nr_device_cfg;

/* ---- context ------------------------------------------------------------- */
/* Replaces the GPU context manager singleton (gpu_context_manager.h:48-101)
 * and the engine factory ISimulationEngine::create (engine_impl.cpp:339-341). */
NR_API int nr_ctx_create(const nr_device_cfg* cfg, nr_ctx** out);
NR_API int nr_ctx_destroy(nr_ctx* ctx);
/* Last error message of this context (or of the library when ctx is NULL). */
NR_API const char* nr_last_error(const nr_ctx* ctx);
NR_API const char* nr_version(void);

/* ---- topology (one-time symbolic analysis, host side) ------------------- */
/* Tensorize (tensors/physics_tensors.cpp:305-493), CSR pattern + stamp map
 * (math/linalg.h:281-465 MatrixConstructor::createCsr/buildPattern), factor
 * graph + Min-Fill (infrastructure/factor_graph.cpp:9-162,
 * infrastructure/treewidth_analyzer.cpp:14-141), symbolic factorisation
 * (math/sparse_lu.cpp:14-128), supernodes and kernel schedules.  `ctx` may
 * be NULL (pure host analysis). */
NR_API int nr_topology_build(nr_ctx* ctx, const nr_netlist* nl, nr_topology** out);
NR_API int nr_topology_destroy(nr_topology* topo);

typedef struct nr_topology_info {
    int32_t n;               /* matrix order (num_nodes)                     */
    int32_t nnz_a;           /* CSR nonzeros of the stamped Jacobian         */
    int32_t nnz_f;           /* nonzeros of the filled L+U pattern           */
    int32_t n_stamps;        /* triplets per iteration (incl. gmin diagonal) */
    int32_t treewidth;       /* Min-Fill estimatedTreewidth                  */
    int32_t ordering_len;    /* raw Min-Fill ordering length (with branch ids)*/
    int32_t fill_in;         /* SymbolicFactorization::estimatedFillIn       */
    int32_t n_supernodes;
    int32_t n_levels;        /* supernodal elimination-tree levels           */
    int32_t max_front;       /* largest frontal matrix order                 */
    int32_t n_res, n_vsrc, n_isrc, n_cap, n_diode, n_mos;
    int64_t lu_flops;        /* multiply-add count x2 of one factorisation   */
    int64_t u_entries;       /* stored factor entries per instance           */
    double analyse_ms;       /* host time of nr_topology_build               */
} nr_topology_info;
NR_API int nr_topology_get_info(const nr_topology* topo, nr_topology_info* info);

/* Integer structures for bit-exact comparison with the reference. */
enum {
    NR_EXPORT_ROW_PTR = 0,       /* int32[n+1]   CSR pattern of A            */
    NR_EXPORT_COL_IDX = 1,       /* int32[nnz_a]                             */
    NR_EXPORT_STAMP_SLOT = 2,    /* int32[n_stamps] triplet -> CSR slot      */
    NR_EXPORT_ORDERING_RAW = 3,  /* int32[ordering_len] Min-Fill ids         */
    NR_EXPORT_PERM = 4,          /* int32[n] perm[new] = old (0-based rows)  */
    NR_EXPORT_FILL_ROW_PTR = 5,  /* int32[n+1] filled L+U pattern            */
    NR_EXPORT_FILL_COL_IDX = 6,  /* int32[nnz_f]                             */
    NR_EXPORT_DIODE_ORDER = 7,   /* int32[n_diode] SoA pos -> netlist index  */
    NR_EXPORT_MOS_ORDER = 8,     /* int32[n_mos]  SoA pos -> netlist index   */
    NR_EXPORT_SN_FIRST = 9       /* int32[n_supernodes] first (permuted)
                                    column of each relaxed supernode of the
                                    multifrontal LU (no reference counterpart) */
};
NR_API int nr_topology_export(const nr_topology* topo, int32_t what,
                              int32_t* out, int64_t capacity, int64_t* written);

/* ---- batches -------------------------------------------------------------- */
/* B instances of one topology, each with its own parameters (MC/corners).
 * Device buffers are allocated here once; solves allocate nothing beyond
 * buffers a first use sizes (the dense LU's workspaces, traces, waveforms).
 * A topology's device tables belong to the context of its first batch:
 * batches of it in another context are refused (NR_ERR_INVALID); build one
 * topology per context.  A topology's launch classes (nr_batch_autotune,
 * nr_batch_set_launch_class) are shared by all its batches, so those calls
 * must not run while another thread solves a batch of the same topology. */
NR_API int nr_batch_alloc(nr_ctx* ctx, const nr_topology* topo, int32_t B,
                          nr_batch** out);
NR_API int nr_batch_free(nr_batch* batch);
NR_API int nr_batch_set_params(nr_batch* batch, int32_t field,
                               const double* values, int64_t count);
/* Seeded start: v0[B * n].  Instances whose seed row is entirely NaN keep the
 * plain seedInitialGuess start (solvers/circuitsim_solver.cpp:32-116). */
NR_API int nr_batch_set_seed(nr_batch* batch, const double* v0, int64_t count);
NR_API int nr_batch_clear_seed(nr_batch* batch);

/* NR_PREC_FP64: the reference's FP64 arithmetic throughout.
 * NR_PREC_FP32_REFINE: device evaluation (the Jacobian entries), assembly and
 * the LU factorisation and solves in FP32; each Newton step solves
 * J dv = b - A v, with the Newton residual b - A v formed in FP64 from the
 * KCL audit vector (the refinement), and updates v += dv in FP64; pnjlim,
 * damping, the KCL audit and the convergence test stay FP64.  Same control
 * flow, iteration accounting and statuses as the FP64 path. */
enum { NR_PREC_FP64 = 0, NR_PREC_FP32_REFINE = 1 };

/* Linear-solve route of each Newton step.  The reference's solveDC solves
 * with the dense partial-pivoting LU solveLU_Pivoted (math/linalg.h:742-854,
 * routed at solvers/circuitsim_solver.cpp:291-295).
 *  NR_LU_AUTO         multifrontal LU on the reference's Min-Fill ordering
 *                     (no pivoting), then each solve's relative residual
 *                     max|b - A x| / max|b| in FP64; a solve above 1e-8,
 *                     not finite or with a pivot below 1e-25 is redone by
 *                     the restated solveLU_Pivoted.  Circuits above
 *                     NR_DENSE_MAX_N nodes take the multifrontal solve
 *                     alone.  On the FP32 path a float solve above 0.1 (or
 *                     not finite, or with a vanishing pivot) is not redone
 *                     in float: the DC solve or transient takes that
 *                     iteration again on the FP64 path from the same
 *                     iterate and finishes there (counted in lu_fallbacks).
 *  NR_LU_MULTIFRONTAL the multifrontal LU alone.
 *  NR_LU_DENSE        the restated solveLU_Pivoted for every solve: the
 *                     reference's LU bit for bit, O(n^3) per Newton step;
 *                     n <= NR_DENSE_MAX_N, else NR_ERR_UNSUPPORTED. */
enum { NR_LU_AUTO = 0, NR_LU_MULTIFRONTAL = 1, NR_LU_DENSE = 2 };
#define NR_DENSE_MAX_N 4096

typedef struct nr_solve_opts {
    int32_t max_iter;        /* per GMIN step, reference MAX_NR_ITER = 50     */
    int32_t gmin_steps;      /* reference GMIN_STEPS = 19                     */
    int32_t precision;       /* NR_PREC_*                                     */
    int32_t seeded_start_step; /* GMIN step where seeded instances start      */
    int32_t record_timing;   /* 1: per-phase HIP event timing; 2: also per
                                kernel launch (nr_batch_get_kernel_stats)    */
    int32_t use_graph;       /* 1: replay each batched iteration as a hipGraph
                                (captured once per grid size and argument
                                set; per-kernel timing is then not kept)   */
    int32_t fallback_plain;  /* 1: an instance whose seeded solve does not
                                converge, or converges with a node beyond
                                1000 x max(5 V, the largest V source) (the
                                KCL audit has no gmin term, so a node only
                                cut-off devices touch passes it at any
                                voltage), is solved again from the plain
                                GMIN-stepping start                           */
    int32_t trace_capacity;  /* > 0: keep each instance's last trace_capacity
                                ConvergenceStep records (nr_batch_get_trace) */
    int32_t lu;              /* NR_LU_*: the linear-solve route              */
    int32_t polish;          /* 0..8: Newton iterations taken after the last
                                GMIN step's convergence test passes (the
                                test accepts KCL < 1e-9 A, which leaves a
                                node error up to ~KCL / G; one more step
                                lands on the root).  They take the full
                                Newton step (no damping: at the root the
                                max-norm KCL no longer falls).  0 is the
                                reference's solveDC; counted in the
                                iterations                                   */
} nr_solve_opts;
NR_API void nr_solve_opts_default(nr_solve_opts* opts);

typedef struct nr_result {
    int32_t B;
    int32_t n_converged;
    int32_t batched_iterations;  /* device-side NR sweeps over the batch      */
    int32_t max_iterations;      /* max per-instance NR iterations           */
    double mean_iterations;
    double max_kcl;              /* max over converged instances             */
    double solve_ms;             /* wall time of nr_batch_solve_dc           */
    double device_evals;         /* diode+MOSFET evaluations performed       */
} nr_result;

/* DC operating point of every instance: solveDC semantics
 * (solvers/circuitsim_solver.cpp:176-555) with GMIN stepping, pnjlim,
 * damping and the KCL test, batched on the GPU. */
NR_API int nr_batch_solve_dc(nr_batch* batch, const nr_solve_opts* opts,
                             nr_result* result);

/* ---- transient ------------------------------------------------------------ */
/* A transient topology also holds the capacitor companion stamps in its
 * pattern, as stepTransient stamps them (solvers/circuitsim_solver.cpp:
 * 559-780, physics/circuitsim.h:2479-2505).  nr_topology_build builds the
 * DC kind. */
enum { NR_TOPO_DC = 0, NR_TOPO_TRANSIENT = 1 };
NR_API int nr_topology_build_ex(nr_ctx* ctx, const nr_netlist* nl, int32_t kind,
                                nr_topology** out);

enum {
    NR_TRAN_START_ZERO = 0,  /* zero voltages, history and arbiter state (the
                                reference's start)                          */
    NR_TRAN_START_SEED = 1   /* the batch seed, e.g. a DC operating point:
                                capacitor history at its branch voltages
                                with zero current, arbiter at that state    */
};
/* Capacitor companion models (IntegrationType, math/integration_method.h:
 * 15-19, same numbering). */
enum { NR_INT_TRAP = 0, NR_INT_GEAR1 = 1, NR_INT_GEAR2 = 2 };
/* Step control. */
enum {
    NR_TRAN_FIXED = 0,       /* `steps` steps of h; a step whose Newton solve
                                fails ends the instance                     */
    NR_TRAN_ARBITER = 1      /* stepTransient + its caller (solvers/
                                circuitsim_solver.cpp:559-780, engine_impl.cpp:
                                205-256): every step starts at h, at most 4
                                attempts halving h on a failed Newton solve or
                                an LTE / energy rejection by the
                                IntegratorArbiter (physics/circuitsim.h:57-368),
                                trapezoidal switched to Gear-2 after 3
                                rejections in a row, until t >= t_stop      */
};
typedef struct nr_transient_opts {
    double h;                /* time step (s); the arbiter's requested step  */
    int32_t steps;           /* NR_TRAN_FIXED: steps to take                 */
    int32_t max_iter;        /* NR iterations per attempt, MAX_NR_ITER = 50  */
    int32_t start;           /* NR_TRAN_START_*                              */
    int32_t n_probe;         /* nodes recorded in the waveform ring          */
    const int32_t* probe;    /* 1-based node ids, n_probe of them            */
    int32_t record_timing;   /* as in nr_solve_opts                          */
    int32_t precision;       /* NR_PREC_*, as in nr_solve_opts               */
    int32_t use_graph;       /* as in nr_solve_opts                          */
    int32_t method;          /* NR_INT_*: the companion model at t = 0       */
    int32_t control;         /* NR_TRAN_*                                    */
    double t_stop;           /* NR_TRAN_ARBITER: stop time (s), <= 1e8 h     */
    int32_t wave_capacity;   /* waveform ring entries per instance; 0: steps+1
                                (fixed) or ceil(t_stop / h) + 1 (arbiter)   */
    int32_t wave_decimation; /* record every k-th accepted step (>= 1), as
                                recordWaveform (shaders/gpu_nr_loop.wgsl:
                                862-891)                                    */
    int32_t lu;              /* NR_LU_*, as in nr_solve_opts                 */
} nr_transient_opts;
NR_API void nr_transient_opts_default(nr_transient_opts* opts);

/* ---- autotuner ------------------------------------------------------------ */
/* No counterpart in the reference (its solves run on the CPU).  Large fronts
 * (order > 8) are factored by one team of lanes per (front, instance); this
 * chooses the team size and threads per block for each bucket of front order
 * by timing the candidate launch shapes on the batch itself (its current
 * parameters; factorisation + back substitution of every instance, `reps`
 * times per candidate, one bucket after another).  The choice is stored in
 * the batch's topology, for the given precision, and used by every later
 * solve of batches of that topology.  It then times the whole-phase kernels
 * (nr_batch_set_persistent) against the tuned per-level launches, and, when
 * those stay, the per-level launches replayed from a hipGraph on one stream
 * against the level fan-out (nr_batch_set_fanout), keeping the faster each
 * time.  Last, on batches of at least 64 instances, it tunes a second class
 * map on graph replays of 32 instances (latency rather than throughput: the
 * tail of a solve, when few instances are left), used for every active count
 * up to the largest power of two where it is still the faster (few_max).
 * It does not change any result: the elimination performs the same
 * operations in the same order whatever the team size and launch mode.
 * Overwrites the batch's solution state; not to be run while a batch of the
 * same topology is solving. */
enum { NR_TUNE_BUCKETS = 8 };
typedef struct nr_autotune_result {
    int32_t n_buckets;
    int32_t bucket_rmax[NR_TUNE_BUCKETS];  /* largest front order in the bucket  */
    int32_t team[NR_TUNE_BUCKETS];         /* chosen lanes per (front, instance) */
    int32_t block[NR_TUNE_BUCKETS];        /* chosen threads per block           */
    int32_t fronts[NR_TUNE_BUCKETS];       /* large fronts in the bucket         */
    int32_t full[NR_TUNE_BUCKETS];         /* 1: the whole front in LDS          */
    double ms_before, ms_after;            /* factor + back substitution of the batch */
    double ms_levels;                      /* the same with the tuned per-level launches */
    double ms_persistent;                  /* the same with the whole-phase kernels (-1: not offered) */
    int32_t persistent;                    /* 1: the whole-phase kernels were kept (faster) */
    double ms_graph;                       /* per-level launches replayed from a hipGraph, one stream (-1: not timed) */
    double ms_graph_fanout;                /* the same with each level's launches on several streams */
    int32_t fanout;                        /* 1: the level fan-out was kept (faster) */
    double ms_few_before, ms_few_after;    /* graph replay of 32 instances with the first map, then
                                              with the classes tuned for them (-1: not tuned) */
    int32_t few_max;                       /* active counts up to this use the second map (0: none) */
} nr_autotune_result;
NR_API int nr_batch_autotune(nr_batch* batch, int32_t precision, int32_t reps,
                             nr_autotune_result* out);
/* Set the launch class of one front-order bucket by hand (a saved tuning
 * table, or a fixed choice): team lanes, threads per block, full = 1 for the
 * whole-front-in-LDS kernel.  NR_ERR_INVALID for a class that does not
 * exist, NR_ERR_UNSUPPORTED when the bucket's fronts would not fit its LDS;
 * like the tuner's choice it is kept by the topology and changes no result. */
NR_API int nr_batch_set_launch_class(nr_batch* batch, int32_t precision, int32_t bucket, int32_t team,
                                     int32_t block, int32_t full);
/* The large-front factorisation and back substitution in one launch each
 * instead of one launch per elimination-tree level and team class: a block
 * per instance walks the levels, its four wavefronts taking the level's
 * fronts in turn, with a block barrier between levels (no dependence between
 * blocks).  It replaces the per-level launches of solveSparse_LU's numeric
 * phase (math/sparse_lu.cpp:134-330) and changes no result: every front gets
 * the same operations in the same order.  on = 0 returns to the per-level
 * launches.  Kept by the topology (every batch of it, as the launch
 * classes); FP64 and FP32 separately (precision NR_PREC_FP64 /
 * NR_PREC_FP32_REFINE); AC and noise keep the per-level launches.
 * NR_ERR_UNSUPPORTED when the largest front's solution values would not fit
 * the back substitution's LDS.  nr_batch_autotune also times it against
 * the tuned per-level launches and keeps the faster. */
NR_API int nr_batch_set_persistent(nr_batch* batch, int32_t precision, int32_t on);
/* Per-level launches of one elimination-tree level on several streams: the
 * team classes' launches of a level (and the small-front classes' launches
 * of a small-front level) go to the context's stream and up to three side
 * streams that fork from it before the level and join it after, so that a
 * level's classes run concurrently; the levels stay in order.  Part of the
 * per-level scheduling of solveSparse_LU's numeric phase
 * (math/sparse_lu.cpp:134-330); no result changes (the fronts of one level
 * are independent).  Off while per-kernel events are recorded
 * (record_timing = 2).  Kept by the topology, FP64 and FP32 separately;
 * AC and noise keep one stream.  nr_batch_autotune times it and keeps the
 * faster.  The library serialises graph captures with its own device
 * allocations and synchronous copies (one process-wide lock); a caller that
 * allocates or copies synchronously from another thread while batches solve
 * with graph replays can still see a capture invalidated (NR_ERR_HIP from
 * the solve), more often with the fan-out on. */
NR_API int nr_batch_set_fanout(nr_batch* batch, int32_t precision, int32_t on);

/* Transient of every instance (the restated stepTransient): capacitor
 * companions of the instance's current method (trapezoidal G = 2C/h with
 * history current, Gear-1 G = C/h, Gear-2 G = 1.5C/h with two voltages of
 * history; math/integration_method.h:47-157), PULSE / SINE sources at
 * t + h, Newton from the arbiter's last voltages (the last accepted step)
 * with the 1e-12 diagonal and no limiting or damping, converged by
 * checkConvergence with the method's capacitor current in the KCL audit,
 * history shifted on acceptance (solvers/circuitsim_solver.cpp:668-699).
 * Instances advance independently: each has its own time, step, method and
 * arbiter state.  The batch's topology must be of kind NR_TOPO_TRANSIENT.
 * In nr_instance_stats, iterations count Newton iterations over all
 * attempts and gmin_steps_done counts accepted steps. */
NR_API int nr_batch_transient(nr_batch* batch, const nr_transient_opts* opts,
                              nr_result* result);
/* Waveform ring of the last transient (per instance, `capacity` entries of a
 * time and n_probe voltages).  Entry k (k = 0 is the start state at t = 0,
 * then every wave_decimation-th accepted step) is stored in slot
 * k mod capacity; slots never written hold NaN.
 *   nr_batch_get_waveform:       out[b][slot][p], count = B*capacity*n_probe
 *   nr_batch_get_waveform_times: out[b][slot],    count = B*capacity
 *   nr_batch_get_wave_recorded:  out[b] = entries written, count = B      */
NR_API int nr_batch_get_waveform(const nr_batch* batch, double* out, int64_t count);
NR_API int nr_batch_get_waveform_times(const nr_batch* batch, double* out, int64_t count);
NR_API int nr_batch_get_wave_recorded(const nr_batch* batch, int64_t* out, int64_t count);
/* Per-instance step control of the last transient. */
typedef struct nr_tran_stats {
    double t_end;            /* time of the last accepted step (s)           */
    double h_last;           /* step of the last attempt (s)                 */
    int32_t accepted;        /* accepted steps                               */
    int32_t rejected;        /* attempts not accepted (Newton failure, LTE /
                                energy rejection, method switch)            */
    int32_t method;          /* NR_INT_* at the end                          */
    int32_t switched_at;     /* accepted steps before the switch to Gear-2, -1
                                if none                                     */
    int32_t instability;     /* the arbiter's instability counter at the end */
    int32_t pad;
} nr_tran_stats;
NR_API int nr_batch_get_tran_stats(const nr_batch* batch, nr_tran_stats* out, int64_t count);

/* ---- AC small-signal analysis -------------------------------------------- */
/* ACSolver::solveAC (solvers/ac_solver.h:339-383) for every instance at its
 * operating point (the batch seed, e.g. the DC solution): Y(jw) = G + jwC
 * with resistors, Norton V-source conductances (1e3 S), capacitors, diode
 * conductances as buildAdmittanceMatrix takes them (:186-196: the derivative
 * of the clamped diode current, at least 1e-15 S), MOSFET gm / gds of the
 * instance's devices (the DC Jacobian's; the reference's own AC builder uses
 * fixed NMOS parameters), the 1e-12 diagonal, and the excitation as
 * solveAC applies it: 1e6 S from the input node to ground with 1e6 x
 * amplitude A injected.  Frequencies: n_points log-spaced from
 * max(f_start, 1e-3) to max(f_stop, f_start + 1) Hz, as solveAC spaces them.
 * Each frequency point is one complex multifrontal factorisation and solve of
 * every instance on the GPU.  The batch's topology must hold the capacitor
 * stamps (NR_TOPO_TRANSIENT), or the circuit have none. */
typedef struct nr_ac_opts {
    double f_start, f_stop;  /* Hz                                          */
    int32_t n_points;        /* frequency points (>= 1)                     */
    int32_t input_node;      /* 1-based excited node; 0: none               */
    double amplitude;        /* excitation (V through the 1e6 S stamp)      */
    int32_t n_probe;         /* 0: every node                               */
    const int32_t* probe;    /* 1-based node ids                            */
    int32_t record_timing;   /* as in nr_solve_opts                         */
    int32_t pad;
} nr_ac_opts;
NR_API void nr_ac_opts_default(nr_ac_opts* opts);
NR_API int nr_batch_ac(nr_batch* batch, const nr_ac_opts* opts, nr_result* result);
/* Results of the last AC run: out[b][f][p][2] (real, imaginary node
 * voltage), count = B * n_points * n_probe * 2; the frequencies, count =
 * n_points; per-instance status (NR_STATUS_SINGULAR if a pivot vanished at
 * some frequency), count = B. */
NR_API int nr_batch_get_ac(const nr_batch* batch, double* out, int64_t count);
NR_API int nr_batch_get_ac_freqs(const nr_batch* batch, double* out, int64_t count);
NR_API int nr_batch_get_ac_status(const nr_batch* batch, int32_t* out, int64_t count);

/* ---- noise analysis ------------------------------------------------------ */
/* ACSolver::solveNoise (solvers/ac_solver.h:524-595) for every instance at
 * its operating point: the output noise PSD S_out(f) = sum_k |H_k|^2 S_k
 * (V^2/Hz) at the output node and sqrt of its trapezoid integral over the
 * frequency points (V rms).  Y(jw) is the AC sweep's without the excitation.
 * Sources in computeNoiseSources order (:400-504): resistors 4kT/R, diodes
 * 2q|I| of the clamped diode current at the operating point, MOSFETs the
 * channel's 4kT (2/3) gm (gm of the instance's device, as in the AC sweep;
 * the reference takes fixed NMOS parameters).  Flicker noise needs KF /
 * C_ox, which the reference defaults to 0 and the netlist does not carry.
 * The reference solves Y V = e_pos - e_neg once per source; here each
 * frequency is one complex factorisation of Y^T and one solve Y^T z = e_out
 * for every instance, and H_k = z[pos] - z[neg] (the same transfer function,
 * rounded differently).  input_node is accepted as solveNoise accepts it and,
 * as there, does not change the result. */
typedef struct nr_noise_opts {
    double f_start, f_stop;  /* Hz, spaced as in nr_ac_opts                 */
    int32_t n_points;        /* frequency points (>= 1)                     */
    int32_t output_node;     /* 1-based observed node; 0: zero noise        */
    double temp_K;           /* temperature of the thermal sources (300.15) */
    int32_t input_node;      /* 1-based, unused (see above)                 */
    int32_t record_timing;   /* as in nr_solve_opts                         */
} nr_noise_opts;
NR_API void nr_noise_opts_default(nr_noise_opts* opts);
NR_API int nr_batch_noise(nr_batch* batch, const nr_noise_opts* opts, nr_result* result);
/* Results of the last noise run: psd[b][f], count = B * n_points; the
 * integrated RMS noise per instance, count = B; the frequencies, count =
 * n_points; per-instance status (NR_STATUS_SINGULAR if a pivot vanished at
 * some frequency), count = B. */
NR_API int nr_batch_get_noise(const nr_batch* batch, double* psd, int64_t count);
NR_API int nr_batch_get_noise_vrms(const nr_batch* batch, double* out, int64_t count);
NR_API int nr_batch_get_noise_freqs(const nr_batch* batch, double* out, int64_t count);
NR_API int nr_batch_get_noise_status(const nr_batch* batch, int32_t* out, int64_t count);

/* Per-instance results.  out arrays are caller-owned host buffers. */
NR_API int nr_batch_get_voltages(const nr_batch* batch, double* out, int64_t count);
enum {
    NR_STATUS_CONVERGED = 0,
    NR_STATUS_NOT_CONVERGED = 1,
    NR_STATUS_SINGULAR = 2
};
typedef struct nr_instance_stats {
    int32_t iterations;          /* true NR iteration count                  */
    int32_t ref_iterations;      /* reference SolverStats accounting         */
    int32_t status;              /* NR_STATUS_*                              */
    int32_t gmin_steps_done;
    double kcl;                  /* final max |KCL| (A), FP64                */
    int32_t seeded;              /* 1: the result comes from the seed; 0: from
                                    the plain start (no seed, or fallback)   */
    int32_t lu_fallbacks;        /* NR_LU_AUTO: Newton steps of the last solve
                                    that the dense pivoting LU solved        */
} nr_instance_stats;
NR_API int nr_batch_get_stats(const nr_batch* batch, nr_instance_stats* out,
                              int64_t count);

/* NR convergence diagnostics of the last DC solve: SolverStats' fields
 * (infrastructure/simtrace.h:76-128) as solveDC fills them
 * (solvers/circuitsim_solver.cpp:407-444, 489-551), per instance.  The
 * failure class is ConvergenceEngine::classifyFailure of the instance's whole
 * history (solvers/convergence_engine.h:26-87).  Not provided: the LU
 * residual and smallest pivot of the reference's pivoting LU (the batched LU
 * does not pivot), the worst-device list and the text fields. */
enum {
    NR_FAIL_NONE = 0, NR_FAIL_SINGULAR = 1, NR_FAIL_DIVERGENCE = 2, NR_FAIL_LIMIT_CYCLE = 3,
    NR_FAIL_SLOW = 4, NR_FAIL_FLOATING_NODE = 5, NR_FAIL_OVERFLOW = 6
};
typedef struct nr_diagnostics {
    int64_t history;             /* ConvergenceStep records (Newton iterations
                                    that reached the update)                 */
    int32_t failure;             /* NR_FAIL_* (ConvergenceFailureType)        */
    int32_t stagnation;          /* stagnationDetected                       */
    int32_t divergence;          /* divergenceDetected                       */
    int32_t worst_node;          /* worstNodeIndex: 1-based, largest |KCL| at
                                    the result (0 if every residual is 0)    */
    double worst_node_v;         /* worstNodeVoltage                          */
    double rms_residual;         /* rmsResidual (A)                          */
    int32_t gmin_stepping_used;  /* gminSteppingUsed: the last GMIN value is
                                    below 1e-3                               */
    int32_t pad;
} nr_diagnostics;
/* One ConvergenceStep (simtrace.h:30-37): residualNorm (max |KCL| after the
 * update), maxVoltageDelta, dampingFactor (0.5 when the damped residual is
 * still not below the last one), worstNode as solveDC fills it (node 1), with
 * the GMIN step and the iteration within it. */
typedef struct nr_conv_step {
    double residual, max_dv, damping;
    int32_t gmin_step, iteration, worst_node, pad;
} nr_conv_step;
NR_API int nr_batch_get_diagnostics(const nr_batch* batch, nr_diagnostics* out, int64_t count);
/* The ring of the last solve with trace_capacity = cap > 0: out[b][k], count
 * = B * cap; record j of instance b (0-based, in the order recorded) is at k
 * = j mod cap, and nr_diagnostics.history gives how many were recorded;
 * slots not reached hold zeros. */
NR_API int nr_batch_get_trace(const nr_batch* batch, nr_conv_step* out, int64_t count);

/* Per-phase device time of the last solve (milliseconds, summed over
 * launches) when record_timing was set. */
enum {
    NR_PHASE_EVAL_STAMP = 0,
    NR_PHASE_FACTOR = 1,
    NR_PHASE_BACKSOLVE = 2,
    NR_PHASE_UPDATE = 3,
    NR_PHASE_COUNT = 4
};
typedef struct nr_timing {
    double ms[NR_PHASE_COUNT];
    int64_t launches[NR_PHASE_COUNT];
    double factor_kernel_ms;     /* factor kernels only                       */
    int64_t factor_instances;    /* instance-factorisations performed         */
} nr_timing;
NR_API int nr_batch_get_timing(const nr_batch* batch, nr_timing* out);

/* Per-kernel accounting of the last solve run with record_timing = 2: one
 * row per kernel symbol, with its launch count, the sum of its per-launch
 * HIP event durations (recorded on the batch's stream around each launch),
 * and the algorithmic HBM bytes and FP64 flops of those launches (sizes of
 * the work itself, independent of how the kernel moves data).  No
 * counterpart in the reference, which has no device kernels on this path;
 * it feeds the roofline figures of the benchmark.  *count receives the
 * number of rows; at most `capacity` are written. */
enum { NR_KERNEL_NAME_LEN = 48 };
typedef struct nr_kernel_stat {
    char name[NR_KERNEL_NAME_LEN];
    int64_t launches;
    double ms;
    double bytes;
    double flops;
} nr_kernel_stat;
NR_API int nr_batch_get_kernel_stats(const nr_batch* batch, nr_kernel_stat* out,
                                     int32_t capacity, int32_t* count);

/* ---- learned seeder behind the NeuralSeeder seam ------------------------- */
/* solvers/neural_seeder.h:45-220 exposes recordSolution / predictSeed /
 * hasHistory / reset around an exponential-average blender.  Here the
 * same seam holds a multilayer perceptron: n_features -> n_hidden layers of
 * `hidden` units with GELU -> n_out node voltages, BF16 weights and
 * activations in the matrix products with FP32 accumulation, FP32 master
 * weights trained with Adam on the GPU.  Features and voltages are
 * normalised internally (per-feature and per-output mean and deviation of
 * the recorded history). */
typedef struct nr_seeder nr_seeder;
typedef struct nr_seeder_cfg {
    int32_t n_features;      /* multiple of 64                               */
    int32_t n_out;           /* node voltages predicted (any size)           */
    int32_t hidden;          /* multiple of 64                               */
    int32_t n_hidden;        /* hidden layers, >= 1                          */
    uint64_t init_seed;      /* weight initialisation                        */
} nr_seeder_cfg;
NR_API int nr_seeder_create(nr_ctx* ctx, const nr_seeder_cfg* cfg, nr_seeder** out);
NR_API int nr_seeder_destroy(nr_seeder* s);
/* recordSolution: count training pairs (features [count x n_features],
 * voltages [count x n_out]) appended to the device-resident history */
NR_API int nr_seeder_record(nr_seeder* s, const float* features, const double* v, int64_t count);
/* hasHistory: number of recorded pairs */
NR_API int64_t nr_seeder_history(const nr_seeder* s);
/* reset: drops the history and re-initialises the weights */
NR_API int nr_seeder_reset(nr_seeder* s);

typedef struct nr_train_opts {
    int32_t epochs;
    int32_t batch;           /* minibatch rows, multiple of 64               */
    float lr, beta1, beta2, eps;
    uint64_t shuffle_seed;
} nr_train_opts;
NR_API void nr_train_opts_default(nr_train_opts* opts);
/* Trains on the recorded history; *loss = mean squared error of the last
 * epoch in normalised units. */
NR_API int nr_seeder_train(nr_seeder* s, const nr_train_opts* opts, float* loss);
/* predictSeed: voltages [count x n_out] for features [count x n_features] */
NR_API int nr_seeder_predict(nr_seeder* s, const float* features, int64_t count, double* v_out);
/* The batch's seed from the seeder, computed on the device (features
 * [B x n_features]); the next nr_batch_solve_dc starts from it at
 * seeded_start_step. */
NR_API int nr_batch_set_seed_from(nr_batch* batch, nr_seeder* s, const float* features);
/* Flat FP32 parameters: per layer W [in x out] row-major, then b [out]. */
NR_API int64_t nr_seeder_param_count(const nr_seeder* s);
NR_API int nr_seeder_get_params(const nr_seeder* s, float* out, int64_t count);
NR_API int nr_seeder_set_params(nr_seeder* s, const float* in, int64_t count);
/* One forward and backward pass without normalisation or update (gradient
 * checks): x [count x n_features], y [count x n_out] in network units,
 * grad = d(mean squared error)/d(params) in the flat layout. */
NR_API int nr_seeder_loss_grad(nr_seeder* s, const float* x, const float* y, int64_t count, float* grad,
                               float* loss);

/* ---- KernelDispatcher slots (tensors/kernel_dispatch.h:24-31) ----------- */
/* Drop-in for batchDiodePhysics_generic (tensors/physics_tensors.cpp:19-51):
 * host arrays in, host arrays out, evaluated on the GPU in FP64. */
NR_API int nr_eval_diodes(nr_ctx* ctx, int32_t count,
                          const int32_t* node_a, const int32_t* node_c,
                          const double* Is, const double* N, const double* Vt,
                          const double* voltages, int32_t n_voltages,
                          double* v_d, double* i_d, double* g_d);
/* Drop-in for batchMosfetPhysics_generic (tensors/physics_tensors.cpp:53-104). */
NR_API int nr_eval_mosfets(nr_ctx* ctx, int32_t count,
                           const int32_t* drains, const int32_t* gates,
                           const int32_t* sources,
                           const double* W, const double* L, const double* Kp,
                           const double* Vth, const double* lambda,
                           const int32_t* isPMOS,
                           const double* voltages, int32_t n_voltages,
                           double* vgs, double* vds, double* ids,
                           double* gm, double* gds);
/* nr_eval_mosfets without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The device copies of the inputs
 * and outputs live in a fixed 8 MiB array on each GPU, held until the process
 * ends: a call may stage at most 8 MiB (4 int32 and 10 double arrays of
 * `count`, plus the voltages, each padded to 256 bytes: about 80,000
 * MOSFETs with as many voltages); larger calls return NR_ERR_INVALID.  Calls
 * on one GPU share that array: the library serialises them, and the results
 * equal nr_eval_mosfets' bit for bit. */
NR_API int nr_eval_mosfets_noalloc(nr_ctx* ctx, int32_t count,
                                   const int32_t* drains, const int32_t* gates,
                                   const int32_t* sources,
                                   const double* W, const double* L, const double* Kp,
                                   const double* Vth, const double* lambda,
                                   const int32_t* isPMOS,
                                   const double* voltages, int32_t n_voltages,
                                   double* vgs, double* vds, double* ids,
                                   double* gm, double* gds);

#ifdef __cplusplus
}
#endif
#endif /* KARAM_NR_H */
