/*
 * oracle.h — CPU FP64 restatement of the reference engine's Newton-Raphson
 * path (DC, transient, AC, noise).  TEST INFRASTRUCTURE ONLY: it is the
 * checker that tests, the smoke entry and the CPU-baseline leg of the
 * benchmark compare the HIP library against.  The product library never
 * links or calls it.
 *
 * Every function cites the reference file:line it restates.  The netlist
 * format is the library's public `nr_netlist` (include/karam_nr.h).
 */
#ifndef KARAM_ORACLE_H
#define KARAM_ORACLE_H

#include <stdint.h>
#include "../include/karam_nr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_soa orc_soa;
typedef struct orc_topo orc_topo;

/* --- a2: tensorizeNetlist ------------------------------------------------ */
orc_soa* orc_tensorize(const nr_netlist* nl);
void orc_soa_free(orc_soa* s);

/* --- a4/a5: generic device kernels --------------------------------------- */
void orc_eval_diodes(int32_t count, const int32_t* a, const int32_t* c,
                     const double* Is, const double* N, const double* Vt,
                     const double* v, int32_t nv,
                     double* vd, double* id, double* gd);
void orc_eval_mosfets(int32_t count, const int32_t* d, const int32_t* g,
                      const int32_t* s, const double* W, const double* L,
                      const double* Kp, const double* Vth, const double* lam,
                      const int32_t* pmos, const double* v, int32_t nv,
                      double* vgs, double* vds, double* ids, double* gm,
                      double* gds);

/* --- a6..a11: topology (pattern, stamp map, Min-Fill, symbolic) ---------- */
int orc_topology_build(const nr_netlist* nl, orc_topo** out);
void orc_topology_free(orc_topo* t);
/* info[0..7] = n, nnz_a, nnz_f, n_stamps, treewidth, ordering_len, fill_in,
 * n_vars */
void orc_topology_info(const orc_topo* t, int32_t* info);
/* `what` uses the NR_EXPORT_* codes of karam_nr.h (plus the two oracle-only
 * codes below). Returns count or -1. */
enum { ORC_EXPORT_STAMP_ROW = 100, ORC_EXPORT_STAMP_COL = 101 };
int64_t orc_topology_export(const orc_topo* t, int32_t what, int32_t* out,
                            int64_t capacity);

/* Triplet values and rhs of one NR iteration (stampSoABlock + gmin), in the
 * reference emission order: vals[n_stamps], rhs[n]. */
int orc_stamp(const nr_netlist* nl, const orc_topo* t, const double* v,
              double gmin, double* vals, double* rhs);
/* CSR values of A from triplet values via the stamp map (assembleFast). */
int orc_assemble(const orc_topo* t, const double* vals, double* a_vals);

/* --- a12: linear solves --------------------------------------------------- */
/* solveLU_Pivoted restated (math/linalg.h:742-854): dense, partial pivot.
 * Returns 0 ok, 2 singular. */
int orc_lu_dense(int32_t n, const int32_t* rp, const int32_t* ci,
                 const double* vals, const double* rhs, double* x,
                 double* smallest_pivot);
/* Same-pattern sparse LU on the symbolic fill pattern, no pivoting. */
int orc_lu_sparse(const orc_topo* t, const double* a_vals, const double* rhs,
                  double* x, double* smallest_pivot);
/* Multifrontal LU on the relaxed supernodes of the same fill pattern,
 * no pivoting, with one fixed operation order (assembly: A entry, then the
 * children's update blocks in child order; elimination: explicit FMAs).
 * The HIP library performs exactly these operations in this order. */
int orc_lu_multifrontal(const orc_topo* t, const double* a_vals, const double* rhs,
                        double* x, double* smallest_pivot);
/* The same in float (FP32 path). */
int orc_lu_multifrontal_f32(const orc_topo* t, const float* a_vals, const float* rhs, float* x);
/* The linear-solve route of the product's default (NR_LU_AUTO): the
 * multifrontal LU, then its relative residual in FP64,
 *   max_i |b_i - (A x)_i| <= bound * max_i |b_i|,
 * each row's residual taken in CSR order.  orc_lu_flag returns 1 when the
 * solve is to be redone by the dense partial-pivoting LU: a value that is not
 * finite, or a residual above the bound (ORC_LU_ETA_FP64 for the FP64 path,
 * ORC_LU_ETA_FP32 for the FP32 path's float system).  *eta (optional): the
 * ratio, or infinity when a value is not finite. */
#define ORC_LU_ETA_FP64 1e-8
#define ORC_LU_ETA_FP32 0.1   /* an inexact Newton step: the FP64 residual corrects it */
int orc_lu_flag(const orc_topo* t, const double* a_vals, const double* rhs, const double* x, double bound,
                double* eta);
/* NR_LU_AUTO on one FP64 system: multifrontal, flagged solves (and singular
 * ones) redone by orc_lu_dense; above NR_DENSE_MAX_N nodes the multifrontal
 * solve alone.  *fallback (optional) is set to 1 when the dense LU gave the
 * result. */
int orc_lu_auto(const orc_topo* t, const double* a_vals, const double* rhs, double* x, int32_t* fallback);
/* FP32 path: the float Jacobian (CSR values) of the DC system at v. */
int orc_jacobian_f32(const nr_netlist* nl, const orc_topo* t, const double* v, double gmin, float* a32);

/* --- a14, a18, a19, a20 ---------------------------------------------------- */
void orc_seed(const nr_netlist* nl, double* v);
double orc_pnjlim(double vnew, double vold, double vt, double vcrit);
double orc_vcrit(double vt, double is);
void orc_apply_pn_limits(const nr_netlist* nl, const orc_topo* t,
                         double* v_new, const double* v_old);
double orc_kcl(const nr_netlist* nl, const orc_topo* t, const double* v,
               double* per_node);
int orc_check_convergence(int32_t n, const double* v_new, const double* v_old,
                          double kcl, double kcl_tol);

/* --- a15: solveDC ---------------------------------------------------------- */
/* ORC_LU_AUTO: orc_lu_auto (the product's default).  On the FP32 path,
 * ORC_LU_DENSE (always) and ORC_LU_AUTO (when orc_lu_flag flags the float
 * multifrontal solve) solve the Newton correction with orc_lu_dense on the
 * FP64 Jacobian and the FP64 Newton residual.  Above NR_DENSE_MAX_N nodes a
 * DC solve or transient on the FP32 path with ORC_LU_AUTO takes a flagged
 * iteration again on the FP64 path and finishes there (one lu_fallback). */
enum { ORC_LU_DENSE = 0, ORC_LU_SPARSE = 1, ORC_LU_MULTIFRONTAL = 2, ORC_LU_AUTO = 3 };
/* ORC_PREC_FP64: the reference's FP64 path.  ORC_PREC_FP32_REFINE: the
 * Jacobian (device evaluation, stamps, assembly) and the multifrontal LU in
 * float; each Newton step solves J dv = b - A v with the Newton residual
 * computed in FP64 and updates v in FP64; limiting, KCL and the convergence
 * test stay FP64. */
enum { ORC_PREC_FP64 = 0, ORC_PREC_FP32_REFINE = 1 };
/* one ConvergenceStep record (infrastructure/simtrace.h:30-37) */
typedef struct orc_conv_step {
    double residual;      /* max |KCL| after the update (A) */
    double max_dv;        /* max |v_new - v_old| (V)        */
    double damping;       /* 0.5 or 1.0, as the reference sets it */
    int32_t gmin_step, iteration, worst_node, pad;
} orc_conv_step;
typedef struct orc_dc_opts {
    int32_t max_iter;     /* 50 */
    int32_t gmin_steps;   /* 19 */
    int32_t lu_kind;      /* ORC_LU_* */
    int32_t start_step;   /* GMIN step to start at (seeded start) */
    int32_t use_seed;     /* 1: v_inout holds the start vector */
    int32_t precision;    /* ORC_PREC_* */
    int32_t trace_cap;    /* > 0: ring of the last trace_cap records */
    int32_t polish;       /* Newton iterations after the last GMIN step
                             converges (the library's nr_solve_opts.polish) */
    orc_conv_step* trace; /* [trace_cap] (per instance in the batch call) */
} orc_dc_opts;
typedef struct orc_dc_out {
    int32_t iterations;      /* true NR iteration count */
    int32_t ref_iterations;  /* reference SolverStats accounting */
    int32_t status;          /* NR_STATUS_* */
    int32_t gmin_steps_done;
    double kcl;
    int32_t iters_per_step[32];
    /* SolverStats diagnostics (infrastructure/simtrace.h:76-128) */
    int64_t history;         /* ConvergenceStep records of the solve */
    int32_t failure;         /* ORC_FAIL_* (classifyFailure)          */
    int32_t stagnation, divergence;
    int32_t worst_node;      /* 1-based, largest |KCL| at the result  */
    int32_t gmin_used;       /* the last GMIN value is below 1e-3     */
    int32_t lu_fallbacks;    /* Newton steps the dense LU solved (ORC_LU_AUTO) */
    double rms;              /* RMS of the per-node KCL at the result */
    double worst_v;          /* the worst node's voltage              */
} orc_dc_out;
/* ConvergenceFailureType (infrastructure/simtrace.h:39-47) */
enum { ORC_FAIL_NONE = 0, ORC_FAIL_SINGULAR = 1, ORC_FAIL_DIVERGENCE = 2, ORC_FAIL_LIMIT_CYCLE = 3,
       ORC_FAIL_SLOW = 4, ORC_FAIL_FLOATING = 5, ORC_FAIL_OVERFLOW = 6 };
void orc_dc_opts_default(orc_dc_opts* o);
int32_t orc_classify_history(const double* residual, const double* max_dv, int64_t n, int32_t singular);
int orc_solve_dc(const nr_netlist* nl, const orc_topo* t,
                 const orc_dc_opts* opts, double* v_inout, orc_dc_out* out);

/* Batch over instances: fields[NR_FIELD_*] point at [B][count] arrays (or
 * NULL for the template value).  OpenMP over instances with `threads`. */
int orc_solve_dc_batch(const nr_netlist* nl, const orc_topo* t, int32_t B,
                       const double* const* fields, const orc_dc_opts* opts,
                       int32_t threads, double* v_out, orc_dc_out* outs);

/* --- a16: transient (stepTransient restated) -------------------------------- */
/* Pattern including the capacitor companion stamps (transient != 0). */
int orc_topology_build_ex(const nr_netlist* nl, int32_t transient, orc_topo** out);
/* The same with the raw Min-Fill ordering given (NR_EXPORT_ORDERING_RAW of an
 * earlier build of the same netlist); treewidth taken from the fill. */
int orc_topology_build_ordered(const nr_netlist* nl, int32_t transient, const int32_t* ordering,
                               int32_t ordering_len, orc_topo** out);
double orc_pulse_value(const double* p, double time);
double orc_sine_value(const double* p, double time);
/* value of a source with waveform kind NR_WAVE_* at `time` (dc: its DC value) */
double orc_wave_value(int32_t kind, const double* par, double dc, double time);
enum { ORC_INT_TRAP = 0, ORC_INT_GEAR1 = 1, ORC_INT_GEAR2 = 2 };   /* IntegrationType */
enum { ORC_TRAN_FIXED = 0, ORC_TRAN_ARBITER = 1 };
typedef struct orc_tran_opts {
    double h;             /* step (s); the arbiter's requested step       */
    int32_t steps;        /* fixed control: steps to take                 */
    int32_t max_iter;     /* 50                                           */
    int32_t start;        /* 0: zero state, 1: v_io holds the state */
    int32_t lu_kind;      /* ORC_LU_*                                     */
    int32_t precision;    /* ORC_PREC_*                                   */
    int32_t method;       /* ORC_INT_* at t = 0                           */
    int32_t control;      /* ORC_TRAN_*                                   */
    double t_stop;        /* arbiter control: stop time (s)               */
    int32_t wave_cap;     /* waveform ring entries (0: default, as in the C ABI) */
    int32_t decimation;   /* record every k-th accepted step (>= 1)       */
} orc_tran_opts;
typedef struct orc_tran_out {
    int32_t iterations;   /* NR iterations over all attempts              */
    int32_t status;       /* NR_STATUS_*                                  */
    int32_t steps_done;   /* accepted steps                               */
    int32_t rejected;     /* attempts not accepted                        */
    int32_t method;       /* ORC_INT_* at the end                         */
    int32_t switched_at;  /* accepted steps before the Gear-2 switch, or -1 */
    int32_t instability;  /* arbiter instability counter at the end       */
    int32_t taa_rejections; /* nodes whose predictor rejectHistory replaced */
    int64_t recorded;     /* waveform entries written                     */
    double kcl;           /* KCL of the last Newton iterate (A)           */
    double t_end;         /* time of the last accepted step               */
    double h_last;        /* step of the last attempt                     */
    int32_t lu_fallbacks; /* Newton steps the dense LU solved (ORC_LU_AUTO) */
    int32_t pad;
} orc_tran_out;
void orc_tran_opts_default(orc_tran_opts* o);
/* resolved waveform ring capacity of `o` */
int32_t orc_wave_capacity(const orc_tran_opts* o);
/* wave_t (optional) [cap], wave (optional) [cap * n_probe]: the waveform
 * ring, entry k in slot k mod cap, NaN where never written */
/* IntegratorArbiter::analyzeStability on a given state (tests): 0 stable
 * (prev / prev2 / prev_energy / prev_dt take the solution), 1 reduce the
 * step, 2 switch the method; *instab in/out, *lte_out = calculateLTE */
int orc_arbiter_analyze(const nr_netlist* nl, const double* cur, const double* hv0, const double* hv1, double dt,
                        double* prev, double* prev2, double* prev_energy, double* prev_dt, int32_t* instab,
                        double* lte_out);
int orc_solve_transient(const nr_netlist* nl, const orc_topo* t, const orc_tran_opts* opts, double* v_io,
                        int32_t n_probe, const int32_t* probe, double* wave_t, double* wave, orc_tran_out* out);
int orc_solve_transient_batch(const nr_netlist* nl, const orc_topo* t, int32_t B, const double* const* fields,
                              const orc_tran_opts* opts, int32_t threads, double* v_io, int32_t n_probe,
                              const int32_t* probe, double* wave_t, double* wave, orc_tran_out* outs);

/* --- AC small-signal analysis (ACSolver::solveAC) ------------------------------ */
typedef struct orc_ac_opts {
    double f_start, f_stop;
    int32_t n_points;
    int32_t input_node;   /* 1-based; 0 none */
    double amplitude;
} orc_ac_opts;
int orc_ac_freqs(const orc_ac_opts* o, double* freqs);
/* out[f][p][2] (re, im); status NR_STATUS_SINGULAR when a pivot vanished */
int orc_solve_ac(const nr_netlist* nl, const orc_topo* t, const double* v_op, const orc_ac_opts* o,
                 int32_t n_probe, const int32_t* probe, double* out, int32_t* status);
int orc_solve_ac_batch(const nr_netlist* nl, const orc_topo* t, int32_t B, const double* const* fields,
                       const double* v_op, const orc_ac_opts* o, int32_t threads, int32_t n_probe,
                       const int32_t* probe, double* out, int32_t* status);

/* --- noise analysis (ACSolver::solveNoise) ------------------------------------ */
enum { ORC_NOISE_PER_SOURCE = 0, ORC_NOISE_ADJOINT = 1 };
typedef struct orc_noise_opts {
    double f_start, f_stop;
    int32_t n_points;
    int32_t output_node;  /* 1-based; 0 gives zero noise */
    double temp_K;
    int32_t mode;         /* ORC_NOISE_*                 */
    int32_t pad;
} orc_noise_opts;
/* psd_out[f] (V^2/Hz), vrms = sqrt of the trapezoid integral */
int orc_solve_noise(const nr_netlist* nl, const orc_topo* t, const double* v_op, const orc_noise_opts* o,
                    double* psd_out, double* vrms, int32_t* status);
int orc_solve_noise_batch(const nr_netlist* nl, const orc_topo* t, int32_t B, const double* const* fields,
                          const double* v_op, const orc_noise_opts* o, int32_t threads, double* psd_out,
                          double* vrms, int32_t* status);

int orc_tensorize_export(const nr_netlist* nl, int32_t* d_a, int32_t* d_c, double* d_is, double* d_n,
                         double* d_vt, int32_t* m_d, int32_t* m_g, int32_t* m_s, int32_t* m_b,
                         double* m_w, double* m_l, double* m_kp, double* m_vth, double* m_lam,
                         int32_t* m_pmos);
double orc_diode_i_clamped(double vd, double Is, double N, double Vt);

#ifdef __cplusplus
}
#endif
#endif
