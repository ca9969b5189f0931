// nr_device.hip — batched Newton-Raphson DC solve on MI355X (gfx950).
//
// One batch = B Monte-Carlo / corner instances of one topology.  One batched
// NR sweep over the still-active instances runs:
//   k_eval          device evaluation, one lane per (device, instance)
//                   (generic kernels, tensors/physics_tensors.cpp:19-104)
//   k_assemble      A and rhs through the precomputed stamp->slot lists, each
//                   slot summed in the reference's stamp order (assembleFast,
//                   math/linalg.h:441-465) — no atomics
//   k_factor_small  multifrontal LU of every small subtree of the supernodal
//                   elimination tree: a wavefront owns a subtree for 64
//                   instances (lane = instance), fronts live in VGPRs and are
//                   assembled through branch-free gather tables
//   k_factor_large  the remaining (large) fronts, level by level: one team per
//                   (front, instance), front staged in LDS, panel factorised
//                   in LDS, trailing update register-tiled and streamed to HBM
//   k_backsolve_*   back substitution, large levels top-down, then the small
//                   subtrees
//   k_update        per instance: unpermute, NaN guard, pnjlim, FP64 KCL
//                   residual, damping, convergence test, GMIN-step control
//                   (solvers/circuitsim_solver.cpp:176-555)
//   k_compact       retire finished instances from the active list
//
// Layout.  Everything the lane-per-instance kernels touch is instance-minor
// ("interleaved": item * Bp + b) so a wavefront's 64 lanes read 512
// contiguous bytes: the small-front pool (their A entries, rhs, update
// blocks and a zero row), small U, y and x.  Everything a per-instance team
// touches is instance-major (b * size + item): the large fronts' A entries,
// rhs, U, y, x and all update blocks consumed by large fronts (including
// those of small subtree roots).  Node voltages and KCL parameters are
// instance-major for the one-block-per-instance update kernel; two small
// transposes bridge the layouts each sweep.
//
// Every floating-point operation of the solve happens in the same order as
// the oracle's restatement (FP contraction off, explicit fma in the LU), so
// GPU and CPU results agree bit for bit wherever the math library calls
// (exp, log) do.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <utility>
#include <utility>
#include <vector>

#include "karam_nr.h"
#include "spice.h"
#include "topology.h"

using karam::Topology;

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static thread_local std::string g_lib_error;

// concurrent launches of one elimination-tree level: the context's stream
// and FAN_WAYS - 1 side streams
constexpr int FAN_WAYS = 4;

struct nr_ctx {
    int device = 0;
    int verbose = 0;
    int cus = 0;                    // compute units of the device
    hipStream_t stream = nullptr;
    // side streams that fork from `stream` and join it again around the
    // launches of one level (default priority, created with the context)
    hipStream_t fan[FAN_WAYS - 1] = {};
    hipEvent_t fan_fork = nullptr;
    hipEvent_t fan_join[FAN_WAYS - 1] = {};
    std::string err;
};

#define HIP_TRY(expr, ctxp)                                                              \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            std::string m_ = std::string(#expr) + ": " + hipGetErrorString(e_);          \
            if (ctxp) (ctxp)->err = m_;                                                  \
            g_lib_error = m_;                                                            \
            return e_ == hipErrorOutOfMemory ? NR_ERR_OOM : NR_ERR_HIP;                  \
        }                                                                                \
    } while (0)

static int set_err(nr_ctx* ctx, int code, const std::string& m) {
    if (ctx) ctx->err = m;
    g_lib_error = m;
    return code;
}

// ---------------------------------------------------------------------------
// device-side topology and batch views (passed by value to kernels)
// ---------------------------------------------------------------------------
// One large front of a launch list: its fields the large-front kernels
// read, in one record (the list entry itself), so that no load depends on a
// supernode index first
struct alignas(16) FrontRec {
    long long u_off, cb_off, rcb_off;
    int c, m, f, a_begin, a_end, ch_begin, ch_end, row_off, J, pad;
};

// One small front of a unit, in the unit's walk order: its fields the
// lane-per-instance kernels read, in one record
struct alignas(16) SmallRec {
    long long u_off, cb_off, rcb_off;
    int c, m, f, nch, tb_off, root, row_off, J;
};

struct alignas(16) ChildRec {
    long long cb_off, rcb_off;   // the child's update block and rhs block
    int nrows, pc, row_off, inv_off;
};

struct DevTopo {
    int n, nnz_a, nD, nM, nR, nV, nI;
    const int *d_a, *d_c;
    const double *d_n, *d_vt, *d_vtraw, *d_pn_vt, *d_pn_isf;
    const int *m_d, *m_g, *m_s, *m_b, *m_pmos;
    const double *m_w, *m_l, *m_lam;
    const int *res_n1, *res_n2, *vsrc_pos, *vsrc_neg, *isrc_pos, *isrc_neg;
    const int* dio_anode_nl;
    const int *lin_ptr, *lin_term, *nl_ptr, *nl_term, *slot_diag;
    const int *rhs_lin_ptr, *rhs_lin_term, *rhs_nl_ptr, *rhs_nl_term;
    const int *kcl_ptr, *kcl_term;
    const int* hub_node;             // nodes with more than KCL_HUB KCL terms
    int n_hub;
    // slots / rhs rows with more than HUB_TERMS device terms (k_assemble_hub)
    const int *hub_slot, *hub_row;
    int n_hub_slot, n_hub_row;
    const int *pn_level_ptr, *pn_diode;
    int n_pn_levels;
    const int *perm, *iperm;
    const int *slot_dst, *node_dst;   // (index << 1) | 1 when the owner front is large
    // supernodes
    const int *sn_first, *sn_ncols, *sn_nrows, *sn_row_off, *sn_a_begin, *sn_a_end;
    const int *sn_child_begin, *sn_child_end, *sn_rclass, *sn_root;
    const long long *sn_u_off, *sn_cb_off, *sn_rcb_off;
    const int *children, *row_idx, *relmap, *fa_slot, *fa_li, *fa_lj;
    // small-front gather tables
    const int *tb_off, *tb;
    // children of large fronts: number of update-block rows that land in the
    // parent's pivot columns (a prefix, the maps are increasing), and the
    // inverse map parent-local row -> child row (or -1), per child
    const int *sn_pc, *inv_off, *inv;
    // per child slot (children[] order): the child's fields the large-front
    // kernels read, in one record so that one wide load, indexed by the slot
    // alone, brings them (no dependent load through children[])
    const struct ChildRec* child_rec;
    // fronts whose panel is kept in HBM: offset of their panel (front-major),
    // for 8-byte values and for the complex path's 16-byte values
    const long long *sn_gp_off, *sn_gp_off_c;
    // the whole-phase large-front kernels: HBM panels (front-major) of the
    // fronts whose panel exceeds a team's LDS share there
    const long long* sn_pgp_off;
    long long zero_item;
    // transient: capacitors and source waveforms (NR_WAVE_*, 7 values each)
    int nC;
    const int *cap_n1, *cap_n2, *vsrc_wave, *isrc_wave;
    const double *vsrc_wave_par, *isrc_wave_par;
    // CSR pattern of A (original node order): the residual check and the
    // dense LU read the assembled values row by row through slot_dst
    const int *row_ptr, *col_idx;
};

// lu_fb: Newton steps of the instance that the dense LU solved (NR_LU_AUTO);
// polished: Newton iterations taken after the last GMIN step converged
struct InstState {
    int step, it, total, ref;
    int status, done, singular, seeded;
    double prev_kcl, kcl;
    int lu_fb, polished;
};

// Per-instance transient step control: the instance's time, the step and
// companion method of the current attempt, and the IntegratorArbiter's
// scalars (physics/circuitsim.h:57-66).  glin_h / glin_method: the step
// and method the linear part of A was last assembled for.
struct TranState {
    double t, h, prev_energy, prev_dt, glin_h;
    long long rec;               // waveform entries written
    int method, attempt, instab, accepted, rejected, switched_at, glin_method, dec;
};
// NR convergence diagnostics of a DC solve (solveDC's ConvergenceStep records
// and SolverStats fields, solvers/circuitsim_solver.cpp:407-444, 489-551).
// classifyFailure (solvers/convergence_engine.h:26-87) reads the whole
// history only through whether some max |dv| exceeded 100 V, the last ten
// residuals, the number of residual direction changes and the length, which
// DiagState keeps record by record; ConvRec is one record (nr_conv_step) of
// the optional per-instance ring; FinalDiag the per-node KCL statistics at
// the result.
struct DiagState {
    double last[10];             // residual of record k at k % 10
    long long n;                 // records
    int changes, big_dv, stagnation, divergence;
};
struct ConvRec {
    double residual, max_dv, damping;
    int gmin_step, iteration, worst_node, pad;
};
struct FinalDiag {
    double rms, worst_v;
    int worst_node, pad;
};

// InstState.done in a transient: 0 solving, 1 finished, 2 the attempt's
// Newton solve converged, 3 it failed (the step control decides), 4 the
// next attempt is armed (k_tran_begin sets it up).  On the FP32 path (DC
// and transient): 5 the instance leaves the FP32 sweeps to finish on FP64
enum { DONE_RUNNING = 0, DONE_FINISHED = 1, DONE_CONVERGED = 2, DONE_FAILED = 3, DONE_ARMED = 4, DONE_PROMOTE = 5 };

struct DevBatch {
    int B, Bp;
    // instance-major parameters (netlist order; Is / KP / VTO in SoA order)
    double *R, *Is, *kp, *vth, *isrc, *vsrc;
    // interleaved parameter copies for the lane-per-instance kernels
    double *IsI, *kpI, *vthI;
    // instance-major node voltages
    double *v, *vn, *xm;
    // interleaved evaluation / assembly state
    double *vI, *glin, *rhslin;
    double *dg, *dieq, *mgds, *mgm, *mieq;
    // small-front pool and factors (interleaved), the large side's A entries
    // and rhs (interleaved) and front blocks (front-major, see fblk); ys / xs
    // hold the forward and backward solutions of every pivot (interleaved)
    double *pool, *Us, *ys, *xs;
    double *Al, *rhsl, *UL, *CBL, *RCBL;
    double* gpanel;              // HBM panels of the fronts too large for LDS
    double* pgpanel;             // the same for the whole-phase kernels (k_factor_persist)
    double* seed;                // optional [B][n]
    // transient: capacitances, capacitor history (branch voltages of the
    // last two accepted steps, current of the last), source values at the
    // attempt's time; vnow / inow = the source values the stamps and the KCL
    // audit use (vsrc / isrc for DC, vsrc_t / isrc_t in a transient); the
    // arbiter's last two stable solutions (instance-major [B][n]); ts = the
    // per-instance step control (NULL for DC)
    double *cap, *hv0, *hv1, *hi0, *vsrc_t, *isrc_t;
    const double *vnow, *inow;
    double *arb_prev, *arb_prev2;
    TranState* ts;
    // per-instance temperature (NR_FIELD_TEMP_K; null: the netlist's ambient
    // for every instance): the diodes' V_T scaled to the instance
    // (interleaved, as the tensorizer scales it) and applyPnJunctionLimits'
    // V_T and Is factor (instance-major)
    double *vtI, *pnvt, *pnisf;
    // FP32 path (p32 = 1): the FP64 KCL vector of the current iterate
    // (instance-major); its interleaved copy lives in rhslin, which the FP32
    // path does not otherwise use
    double* res;
    int p32;
    // per-instance MOSFET drain currents of the KCL audit in progress
    // (instance-major [B][nM]; null: evaluated per KCL term)
    double* mos_i;
    // DC diagnostics: per-instance state, the optional ring of the last
    // trace_cap records per instance, the final per-node statistics
    DiagState* diag;
    ConvRec* trace;
    int trace_cap;
    FinalDiag* fin;
    InstState* st;
    int* act;
    int* nact;
    // linear-solve route (NR_LU_*).  AUTO: per instance, the bits of
    // max |b - A x| and max |b| (non-negative doubles order as their bits)
    // and a non-finite flag, gathered by k_linres and cleared by k_lu_flag;
    // the list of instances the dense LU solves, with dn_ctl[0] = their count
    // and dn_ctl[1] = the head of the queue its workgroups take them from;
    // dn_slots workspaces of n * n doubles (null when the route does not use
    // the dense LU)
    int lu;
    // FP32 path above NR_DENSE_MAX_N nodes with the AUTO route (DC and
    // transient): an instance whose float solve k_lu_flag rejects (or whose
    // float LU met a vanishing pivot) drops that iteration and finishes on
    // the FP64 path from the same iterate
    int promote;
    // DC: Newton iterations to take after the last GMIN step converges
    // (nr_solve_opts.polish)
    int polish;
    int dn_slots;
    unsigned long long* dn_acc;
    int* dn_bad;
    int* dn_list;
    int* dn_ctl;
    double* dn_mat;
};

__device__ __forceinline__ size_t IX(long long item, int b, int Bp) { return (size_t)item * (size_t)Bp + (size_t)b; }

// Complex double of the AC path.  Each operation is written out with its
// rounding steps (explicit fma, no contraction) so that the oracle, which
// uses the same formulas, gives bit-identical results:
//   a + b = (ar + br, ai + bi);  fma(a, b, c) = a b + c with
//   re = fma(ar, br, fma(-ai, bi, cr)), im = fma(ar, bi, fma(ai, br, ci));
//   a / b = ((fma(ar, br, ai bi)) / d, fma(ai, br, -(ar bi)) / d), d = fma(br, br, bi bi)
struct cdouble {
    double re, im;
    __host__ __device__ constexpr cdouble(double r = 0.0, double i = 0.0) : re(r), im(i) {}
};
__device__ __forceinline__ cdouble operator+(cdouble a, cdouble b) { return cdouble(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cdouble& operator+=(cdouble& a, cdouble b) { a = a + b; return a; }
__device__ __forceinline__ cdouble operator-(cdouble a) { return cdouble(-a.re, -a.im); }
__device__ __forceinline__ cdouble vfma(cdouble a, cdouble b, cdouble c) {
    return cdouble(__builtin_fma(a.re, b.re, __builtin_fma(-a.im, b.im, c.re)),
                   __builtin_fma(a.re, b.im, __builtin_fma(a.im, b.re, c.im)));
}
__device__ __forceinline__ cdouble operator/(cdouble a, cdouble b) {
#pragma clang fp contract(off)
    const double d = __builtin_fma(b.re, b.re, b.im * b.im);
    return cdouble(__builtin_fma(a.re, b.re, a.im * b.im) / d, __builtin_fma(a.im, b.re, -(a.re * b.im)) / d);
}
__device__ __forceinline__ double cmag(cdouble a) {
    const double x = __builtin_fabs(a.re), y = __builtin_fabs(a.im);
    return x >= y ? x : y;
}

// ---------------------------------------------------------------------------
// device helpers (reference arithmetic; FP contraction off)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double volt(const double* v, int n, int node) {
    return (node > 0 && node <= n) ? v[node - 1] : 0.0;
}

// V_T of diode i in instance b: the tensorizer's V_T * T / 300.15
// (tensors/physics_tensors.cpp:419) at the instance's temperature, or the
// netlist's ambient
__device__ __forceinline__ double diode_vt(const DevTopo& T, const DevBatch& P, int i, int b) {
    return P.vtI ? P.vtI[IX(i, b, P.Bp)] : T.d_vt[i];
}

// batchDiodePhysics_generic (tensors/physics_tensors.cpp:19-51)
__device__ __forceinline__ void diode_eval(double vd, double Is, double N, double Vt, double& i, double& g) {
#pragma clang fp contract(off)
    double n_vt = N * Vt;
    double vcrit = 30.0 * n_vt;
    double arg = vd / n_vt;
    double argc = (30.0 < arg) ? 30.0 : arg;          // std::min(arg, 30.0)
    double e = exp(argc);
    double ib = Is * (e - 1.0);
    double gb = (Is / n_vt) * e;
    double d = vd - vcrit;
    double delta = (0.0 < d) ? d : 0.0;               // std::max(0.0, d)
    i = ib + delta * gb;
    g = (1e-12 < gb) ? gb : 1e-12;                    // std::max(GMIN, gb)
}

// diode_model_pair current (physics/device_physics.h:72-102), KCL audit
__device__ __forceinline__ double diode_i_clamped(double vd, double Is, double N, double Vt) {
#pragma clang fp contract(off)
    double n_vt = N * Vt;
    double arg = vd / n_vt;
    if (arg > 30.0) {
        double em = exp(30.0);
        double imax = Is * (em - 1.0);
        double gmax = (Is / n_vt) * em;
        double delta = vd - (n_vt * 30.0);
        return imax + delta * gmax;
    }
    double ev = exp(arg);
    return Is * (ev - 1.0);
}

// batchMosfetPhysics_generic (tensors/physics_tensors.cpp:53-104)
__device__ __forceinline__ void mos_eval(double vg, double vd_, double vs, double W, double L, double Kp,
                                         double Vth, double lam, int pmos, double& vgs, double& vds,
                                         double& ids, double& gm, double& gds) {
#pragma clang fp contract(off)
    double sign = pmos ? -1.0 : 1.0;
    vgs = vg - vs;
    vds = vd_ - vs;
    double lvgs = sign * vgs, lvds = sign * vds, lvth = sign * Vth;
    double K = 0.5 * Kp * (W / L);
    double vov = lvgs - lvth;
    if (vov <= 0.0) {
        ids = 0.0; gm = 0.0; gds = 1e-12;
    } else if (lvds < vov) {
        ids = sign * K * (2.0 * vov * lvds - lvds * lvds);
        gm = 2.0 * K * lvds;
        gds = 2.0 * K * (vov - lvds);
    } else {
        ids = sign * K * vov * vov * (1.0 + lam * lvds);
        gm = 2.0 * K * vov * (1.0 + lam * lvds);
        gds = K * vov * vov * lam;
    }
    if (gds < 1e-12) gds = 1e-12;
}

// pnjlim (physics/device_physics.h:132-168)
__device__ __forceinline__ double pnjlim(double vnew, double vold, double vt, double vcrit) {
#pragma clang fp contract(off)
    double vlim = vnew;
    if (vold > 30.0 * vt && vnew > 30.0 * vt) return vnew;
    if (vold < -5.0 * vt && vnew < -5.0 * vt) return vnew;
    if (vnew > vcrit && fabs(vnew - vold) > 2.0 * vt) {
        if (vold > 0.0) {
            double arg = 1.0 + (vnew - vold) / vt;
            if (arg > 0.0) vlim = vold + vt * log(arg);
            else vlim = vcrit;
        } else {
            vlim = vt * log(vnew / vt);
        }
    } else if (vnew < -5.0 * vt) {
        double lo = vold - 5.0 * vt;
        vlim = (vnew < lo) ? lo : vnew;               // std::max(vnew, lo)
    }
    return vlim;
}

__device__ __forceinline__ int term_kind(int t) { return (t >> 1) & 7; }
__device__ __forceinline__ bool term_neg(int t) { return (t & 1) != 0; }
__device__ __forceinline__ int term_idx(int t) { return t >> 4; }

// std::max(a, b) semantics: a NaN candidate never replaces the running value
__device__ __forceinline__ double ref_max(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ double block_max(double x, double* red) {
    for (int o = 32; o > 0; o >>= 1) x = ref_max(x, __shfl_xor(x, o));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = x;
    __syncthreads();
    double r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = ref_max(r, red[i]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ int block_and(int x, int* red) {
    x = __all(x) ? 1 : 0;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = x;
    __syncthreads();
    int r = 1;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r &= red[i];
    __syncthreads();
    return r;
}

// KCL nodes with more terms than this (a supply rail feeding every PMOS
// source) are summed by the whole block, HUB_CHUNK terms at a time
constexpr int KCL_HUB = 256;
constexpr int HUB_CHUNK = 1024;

// One KCL term (calculatePhysicalResiduals, physics/circuitsim.h:1250-1671,
// MOSFET term restated with SoA parameters), with its sign applied
// mi (optional): the instance's MOSFET drain currents at v, evaluated once
// per audit (block_mos_currents); a MOSFET term then reads its current there
__device__ double kcl_term_val(const DevTopo& T, const DevBatch& P, int b, const double* v, int t,
                               const double* mi = nullptr) {
#pragma clang fp contract(off)
    const int n = T.n;
    const int idx = term_idx(t);
    double val;
    switch (term_kind(t)) {
    case karam::TK_RES:
        val = (volt(v, n, T.res_n1[idx]) - volt(v, n, T.res_n2[idx])) / P.R[(size_t)b * T.nR + idx];
        break;
    case karam::TK_VSRC:
        val = (P.vnow[(size_t)b * T.nV + idx] - (volt(v, n, T.vsrc_pos[idx]) - volt(v, n, T.vsrc_neg[idx]))) /
              1e-3;
        break;
    case karam::TK_CAP: {   // displacement current of the attempt's companion (:1460-1490)
        const size_t ci = (size_t)b * T.nC + idx;
        const double vc = volt(v, n, T.cap_n1[idx]) - volt(v, n, T.cap_n2[idx]);
        const double C = P.cap[ci], h = P.ts[b].h;
        const int m = P.ts[b].method;
        if (m == NR_INT_TRAP) val = ((2.0 * C) / h) * (vc - P.hv0[ci]) - P.hi0[ci];
        else if (m == NR_INT_GEAR2) val = (C / h) * (1.5 * vc - 2.0 * P.hv0[ci] + 0.5 * P.hv1[ci]);
        else val = (C / h) * (vc - P.hv0[ci]);
        break;
    }
    case karam::TK_ISRC:
        val = P.inow[(size_t)b * T.nI + idx];
        break;
    case karam::TK_DIODE: {
        // at the V_T the stamps use: the reference's audit takes the
        // unscaled V_T (physics/circuitsim.h:1346-1347) and so can never pass
        // away from 300.15 K, where the scale is exactly 1
        double vd = volt(v, n, T.d_a[idx]) - volt(v, n, T.d_c[idx]);
        val = diode_i_clamped(vd, P.Is[(size_t)b * T.nD + idx], T.d_n[idx], diode_vt(T, P, idx, b));
        break;
    }
    default: {  // TK_MOS
        if (mi) {
            val = mi[idx];
            break;
        }
        double vgs, vds, ids, gm, gds;
        mos_eval(volt(v, n, T.m_g[idx]), volt(v, n, T.m_d[idx]), volt(v, n, T.m_s[idx]), T.m_w[idx],
                 T.m_l[idx], P.kp[(size_t)b * T.nM + idx], P.vth[(size_t)b * T.nM + idx], T.m_lam[idx],
                 T.m_pmos[idx], vgs, vds, ids, gm, gds);
        val = ids;
        break;
    }
    }
    return term_neg(t) ? -val : val;
}

// KCL residual of one node: its terms summed in list order
__device__ double kcl_node(const DevTopo& T, const DevBatch& P, int b, const double* v, int i,
                           const double* mi = nullptr) {
#pragma clang fp contract(off)
    double s = 0.0;
    for (int k = T.kcl_ptr[i]; k < T.kcl_ptr[i + 1]; ++k) s = s + kcl_term_val(T, P, b, v, T.kcl_term[k], mi);
    return s;
}

// the MOSFET drain currents of instance b at v, one evaluation per device
// (the KCL list holds two terms per MOSFET), into mi; whole block
__device__ void block_mos_currents(const DevTopo& T, const DevBatch& P, int b, const double* v, double* mi) {
#pragma clang fp contract(off)
    const int n = T.n;
    for (int k = threadIdx.x; k < T.nM; k += blockDim.x) {
        double vgs, vds, ids, gm, gds;
        mos_eval(volt(v, n, T.m_g[k]), volt(v, n, T.m_d[k]), volt(v, n, T.m_s[k]), T.m_w[k], T.m_l[k],
                 P.kp[(size_t)b * T.nM + k], P.vth[(size_t)b * T.nM + k], T.m_lam[k], T.m_pmos[k], vgs, vds, ids,
                 gm, gds);
        mi[k] = ids;
    }
    __syncthreads();
}

// max |KCL| over the nodes of instance b (whole block); put(i, r) receives
// every node's value.  A hub node's terms are evaluated by all threads into
// LDS and added by thread 0 in list order, so each node's value is the same
// sequential sum as kcl_node's; the maximum does not depend on the order.
template <typename Put>
__device__ double block_kcl_put(const DevTopo& T, const DevBatch& P, int b, const double* v, double* red, Put put) {
#pragma clang fp contract(off)
    __shared__ double hbuf[HUB_CHUNK];
    double* mi = (P.mos_i && T.nM > 0) ? P.mos_i + (size_t)b * T.nM : nullptr;
    if (mi) block_mos_currents(T, P, b, v, mi);
    double mx = 0.0;
    for (int i = threadIdx.x; i < T.n; i += blockDim.x) {
        if (T.kcl_ptr[i + 1] - T.kcl_ptr[i] > KCL_HUB) continue;
        const double r = kcl_node(T, P, b, v, i, mi);
        put(i, r);
        mx = ref_max(mx, fabs(r));
    }
    for (int h = 0; h < T.n_hub; ++h) {
        const int i = T.hub_node[h];
        const int k0 = T.kcl_ptr[i], k1 = T.kcl_ptr[i + 1];
        double s = 0.0;
        for (int base = k0; base < k1; base += HUB_CHUNK) {
            const int cnt = min(HUB_CHUNK, k1 - base);
            __syncthreads();
            for (int q = threadIdx.x; q < cnt; q += blockDim.x)
                hbuf[q] = kcl_term_val(T, P, b, v, T.kcl_term[base + q], mi);
            __syncthreads();
            if (threadIdx.x == 0)
                for (int q = 0; q < cnt; ++q) s = s + hbuf[q];
        }
        if (threadIdx.x == 0) {
            put(i, s);
            mx = ref_max(mx, fabs(s));
        }
    }
    return block_max(mx, red);
}

// max |KCL| over the nodes of instance b; `out` (optional) receives the
// per-node values
__device__ double block_kcl(const DevTopo& T, const DevBatch& P, int b, const double* v, double* red,
                            double* out = nullptr) {
    return block_kcl_put(T, P, b, v, red, [&](int i, double r) { if (out) out[i] = r; });
}

// FP32 path: the interleaved KCL vector at v (whole block per instance)
__device__ void residual_to_interleaved(const DevTopo& T, const DevBatch& P, int b, const double* v) {
    __shared__ double red[8];
    (void)block_kcl_put(T, P, b, v, red, [&](int i, double r) { P.rhslin[IX(i, b, P.Bp)] = r; });
}

// Capacitor companion of the attempt's method (math/integration_method.h:
// 47-157): its conductance, and the history current it injects into its
// first plate (trapezoidal -I_eq = G v0 + i0, Gear-1 G v0, Gear-2
// C (2/h v0 - 0.5/h v1)), each written as the reference computes it
__device__ double cap_conductance(const DevBatch& P, int b, int nC, int idx) {
#pragma clang fp contract(off)
    const double C = P.cap[(size_t)b * nC + idx], h = P.ts[b].h;
    const int m = P.ts[b].method;
    return m == NR_INT_TRAP ? (2.0 * C) / h : m == NR_INT_GEAR2 ? (1.5 * C) / h : C / h;
}

__device__ double cap_history_current(const DevBatch& P, int b, int nC, int idx) {
#pragma clang fp contract(off)
    const size_t ci = (size_t)b * nC + idx;
    const double C = P.cap[ci], h = P.ts[b].h;
    switch (P.ts[b].method) {
    case NR_INT_TRAP: {
        const double G = (2.0 * C) / h;
        return -(-G * P.hv0[ci] - P.hi0[ci]);
    }
    case NR_INT_GEAR2: return C * ((2.0 / h) * P.hv0[ci] - (0.5 / h) * P.hv1[ci]);
    default: {
        const double G = C / h;
        return -(-G * P.hv0[ci]);
    }
    }
}

// linear rhs of node i in emission order: Norton V sources (at the values in
// P.vnow), I sources (P.inow), then the capacitor history currents
__device__ double rhs_lin_node(const DevTopo& T, const DevBatch& P, int b, int i) {
#pragma clang fp contract(off)
    double r = 0.0;
    for (int k = T.rhs_lin_ptr[i]; k < T.rhs_lin_ptr[i + 1]; ++k) {
        const int t = T.rhs_lin_term[k];
        const int idx = term_idx(t);
        double val;
        switch (term_kind(t)) {
        case karam::TK_VSRC: val = P.vnow[(size_t)b * T.nV + idx] * 1.0e3; break;
        case karam::TK_CAP: val = cap_history_current(P, b, T.nC, idx); break;
        default: val = P.inow[(size_t)b * T.nI + idx]; break;
        }
        r = term_neg(t) ? r + (-val) : r + val;
    }
    return r;
}

// PULSE source value at time t (physics/circuitsim.h:2155-2170);
// p = v1 v2 td tr tf pw per
__device__ double pulse_value(const double* p, double time) {
#pragma clang fp contract(off)
    const double v1 = p[0], v2 = p[1], td = p[2], tr = p[3], tf = p[4], pw = p[5], per0 = p[6];
    double t = time - td;
    if (!(t > 0)) return v1;
    const double per = per0 > 0 ? per0 : 1e99;
    t = fmod(t, per);
    if (t < tr) return v1 + (v2 - v1) * (t / tr);
    if (t < tr + pw) return v2;
    if (t < tr + pw + tf) return v2 + (v1 - v2) * ((t - tr - pw) / tf);
    return v1;
}

// SINE source value at time t (physics/circuitsim.h:2171-2180);
// p = vo va freq td theta phase_deg
__device__ double sine_value(const double* p, double time) {
#pragma clang fp contract(off)
    const double vo = p[0], va = p[1], freq = p[2], td = p[3], theta = p[4], ph = p[5];
    const double t = time - td;
    if (t > 0) {
        const double omega = 2.0 * 3.141592653589793 * freq;
        const double damping = exp(-t * theta);
        const double phase = ph * (3.141592653589793 / 180.0);
        return vo + va * sin(omega * t + phase) * damping;
    }
    return vo;
}

__device__ double wave_value(int kind, const double* par, double dc, double time) {
    return kind == NR_WAVE_PULSE ? pulse_value(par, time) : kind == NR_WAVE_SINE ? sine_value(par, time) : dc;
}

// Sum of f(0..count-1) added in index order from 0.0 (the reference's
// sequential accumulation), by the whole block: terms are evaluated in
// parallel into LDS, thread 0 adds them; every thread gets the sum.
template <typename F>
__device__ double block_ordered_sum(int count, F f) {
#pragma clang fp contract(off)
    __shared__ double buf[HUB_CHUNK];
    __shared__ double out;
    double s = 0.0;
    for (int base = 0; base < count; base += HUB_CHUNK) {
        const int cnt = min(HUB_CHUNK, count - base);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt; q += blockDim.x) buf[q] = f(base + q);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < cnt; ++q) s = s + buf[q];
    }
    if (threadIdx.x == 0) out = s;
    __syncthreads();
    const double r = out;
    __syncthreads();
    return r;
}

// stored capacitor energy sum 0.5 C v^2 in netlist order (analyzeStability,
// physics/circuitsim.h:289-295)
__device__ double cap_energy(const DevTopo& T, const DevBatch& P, int b, const double* v) {
    const int n = T.n;
    return block_ordered_sum(T.nC, [&](int i) {
#pragma clang fp contract(off)
        double x = 0.0;
        const int p = T.cap_n1[i], q = T.cap_n2[i];
        if (p > 0 && p <= n) x += v[p - 1];
        if (q > 0 && q <= n) x -= v[q - 1];
        return 0.5 * P.cap[(size_t)b * T.nC + i] * x * x;
    });
}

// linear part of A (resistors, Norton conductances, capacitor companions of
// the instance's current step and method) into glin, by the whole block
__device__ void lin_matrix(const DevTopo& T, const DevBatch& P, int b) {
#pragma clang fp contract(off)
    for (int s = threadIdx.x; s < T.nnz_a; s += blockDim.x) {
        double a = 0.0;
        for (int k = T.lin_ptr[s]; k < T.lin_ptr[s + 1]; ++k) {
            const int t = T.lin_term[k];
            const int idx = term_idx(t);
            double g;
            switch (term_kind(t)) {
            case karam::SK_RES: g = 1.0 / P.R[(size_t)b * T.nR + idx]; break;
            case karam::SK_CAP: g = cap_conductance(P, b, T.nC, idx); break;
            default: g = 1.0e3; break;
            }
            a = term_neg(t) ? a + (-g) : a + g;
        }
        P.glin[IX(s, b, P.Bp)] = a;
    }
}

// ---------------------------------------------------------------------------
// kernels: set-up, evaluation, assembly
// ---------------------------------------------------------------------------

// seedInitialGuess (solvers/circuitsim_solver.cpp:32-116) into v, by a whole
// block: V+ nodes at their source value, diode anodes at least 0.6 V, then
// unset I-source + nodes at half the supply
__device__ void seed_initial_guess(const DevTopo& T, const DevBatch& P, int b, double* v) {
    const int n = T.n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double* V = P.vsrc + (size_t)b * T.nV;
        for (int i = 0; i < T.nV; ++i) {
            int p = T.vsrc_pos[i];
            if (p > 0 && p <= n) v[p - 1] = V[i];
        }
        for (int i = 0; i < T.nD; ++i) {   // netlist order
            int a = T.dio_anode_nl[i];
            if (a > 0 && a <= n) v[a - 1] = ref_max(v[a - 1], 0.6);
        }
        double vs = 5.0;
        for (int i = 0; i < T.nV; ++i)
            if (V[i] > vs) vs = V[i];
        for (int i = 0; i < T.nI; ++i) {
            int p = T.isrc_pos[i];
            if (p <= 0 || p > n) continue;
            if (v[p - 1] == 0.0) v[p - 1] = vs * 0.5;
        }
    }
    __syncthreads();
}

// Plain-start fallback (one block per instance): an instance whose seeded
// solve did not converge restarts from seedInitialGuess at the first GMIN
// step; its iteration counts keep accumulating.
// A seeded instance restarts from the plain start when its seeded solve did
// not converge, or converged with a node beyond vbound: the KCL audit holds
// no gmin term, so a node that only cut-off devices touch meets it at any
// voltage, and a seed can leave one there (the plain start's GMIN stepping
// does not).
__global__ __launch_bounds__(256) void k_fallback(DevTopo T, DevBatch P, double vbound) {
    __shared__ double red[4];
    const int b = blockIdx.x;
    InstState st = P.st[b];
    double* v = P.v + (size_t)b * T.n;
    if (st.status == NR_STATUS_CONVERGED) {
        double mx = 0.0;
        for (int i = threadIdx.x; i < T.n; i += blockDim.x) {
            const double a = v[i] < 0.0 ? -v[i] : v[i];
            mx = (mx < a || a != a) ? a : mx;
        }
        mx = block_max(mx, red);
        if (!st.seeded || mx <= vbound) return;
    }
    seed_initial_guess(T, P, b, v);
    for (int i = threadIdx.x; i < T.n; i += blockDim.x) P.vI[IX(i, b, P.Bp)] = v[i];
    if (P.p32) residual_to_interleaved(T, P, b, v);
    if (threadIdx.x == 0) {
        st.step = 0; st.it = 0;
        st.status = NR_STATUS_NOT_CONVERGED; st.done = 0; st.singular = 0; st.seeded = 0;
        st.prev_kcl = 0.0; st.kcl = 0.0; st.polished = 0;
        P.st[b] = st;
    }
}

// Start state of one instance (one block per instance) and the per-instance
// linear part of A and rhs.  mode 0 (DC): seedInitialGuess
// (solvers/circuitsim_solver.cpp:32-116) or the caller's seed; mode 1
// (transient): zero voltages; mode 2 (transient): the caller's seed.  A
// transient also starts its step control (time 0, step h, method), its
// capacitor history (branch voltages of the start state, zero current) and
// its arbiter (the start state, prev_dt = 1e-9, energy 0 from the zero
// start as IntegratorArbiter::initialize sets it, physics/circuitsim.h:
// 135-173, or the seed's stored energy); the rhs is set per attempt.
__global__ __launch_bounds__(256) void k_init(DevTopo T, DevBatch P, int seeded_step, int mode, double h,
                                              int method) {
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    const int n = T.n, Bp = P.Bp;
    double* v = P.v + (size_t)b * n;
    __shared__ int use_seed;
    if (threadIdx.x == 0) {
        int ok = 0;
        if (P.seed && mode != 1) {
            const double* sv = P.seed + (size_t)b * n;
            for (int i = 0; i < n; ++i)
                if (sv[i] == sv[i]) { ok = 1; break; }
        }
        use_seed = ok;
    }
    __syncthreads();
    if (use_seed) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = P.seed[(size_t)b * n + i];
    } else if (mode != 0) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = 0.0;
    } else {
        seed_initial_guess(T, P, b, v);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) P.vI[IX(i, b, Bp)] = v[i];
    if (mode != 0) {
        for (int i = threadIdx.x; i < T.nC; i += blockDim.x) {
            const size_t ci = (size_t)b * T.nC + i;
            P.hv0[ci] = volt(v, n, T.cap_n1[i]) - volt(v, n, T.cap_n2[i]);
            P.hv1[ci] = P.hv0[ci];
            P.hi0[ci] = 0.0;
        }
        if (P.arb_prev)
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                P.arb_prev[(size_t)b * n + i] = v[i];
                P.arb_prev2[(size_t)b * n + i] = v[i];
            }
        const double e0 = use_seed ? cap_energy(T, P, b, v) : 0.0;
        if (threadIdx.x == 0) {
            TranState ts;
            ts.t = 0.0; ts.h = h; ts.prev_energy = e0; ts.prev_dt = 1e-9; ts.glin_h = h;
            ts.rec = 0; ts.method = method; ts.attempt = 0; ts.instab = 0; ts.accepted = 0; ts.rejected = 0;
            ts.switched_at = -1; ts.glin_method = method; ts.dec = 0;
            P.ts[b] = ts;
        }
        __syncthreads();
    }
    lin_matrix(T, P, b);
    if (mode == 0 && P.p32) residual_to_interleaved(T, P, b, v);
    else if (mode == 0)
        for (int i = threadIdx.x; i < n; i += blockDim.x) P.rhslin[IX(i, b, Bp)] = rhs_lin_node(T, P, b, i);
    if (threadIdx.x == 0) {
        InstState s;
        s.step = use_seed ? seeded_step : 0;
        s.it = 0; s.total = 0; s.ref = 0;
        s.status = NR_STATUS_NOT_CONVERGED; s.done = mode != 0 ? DONE_ARMED : DONE_RUNNING; s.singular = 0;
        s.seeded = use_seed;
        s.prev_kcl = 0.0; s.kcl = 0.0;
        s.lu_fb = 0; s.polished = 0;
        P.st[b] = s;
        if (mode == 0) {
            DiagState d;
            for (int k = 0; k < 10; ++k) d.last[k] = 0.0;
            d.n = 0; d.changes = 0; d.big_dv = 0; d.stagnation = 0; d.divergence = 0;
            P.diag[b] = d;
        }
        P.act[b] = b;
        if (b == 0) *P.nact = P.B;
    }
}

// Transient attempt set-up (one block per armed instance of the active
// list): source values at t + h (PULSE / SINE sources from their waveform,
// the others at their DC value), the linear part of A when the step or the
// method changed, the interleaved copy of the Newton estimate, and the
// linear rhs with the capacitor history (FP32 path: the KCL vector at the
// estimate).  The instance then runs.
__global__ __launch_bounds__(256) void k_tran_begin(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                    int nact_grid) {
#pragma clang fp contract(off)
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    if ((int)blockIdx.x >= nact) return;
    const int b = act[blockIdx.x];
    if (P.st[b].done != DONE_ARMED) return;
    const TranState ts = P.ts[b];
    const double tn = ts.t + ts.h;
    for (int i = threadIdx.x; i < T.nV; i += blockDim.x)
        P.vsrc_t[(size_t)b * T.nV + i] =
            wave_value(T.vsrc_wave[i], T.vsrc_wave_par + 7 * i, P.vsrc[(size_t)b * T.nV + i], tn);
    for (int i = threadIdx.x; i < T.nI; i += blockDim.x)
        P.isrc_t[(size_t)b * T.nI + i] =
            wave_value(T.isrc_wave[i], T.isrc_wave_par + 7 * i, P.isrc[(size_t)b * T.nI + i], tn);
    if (ts.glin_h != ts.h || ts.glin_method != ts.method) {
        lin_matrix(T, P, b);
        if (threadIdx.x == 0) {
            P.ts[b].glin_h = ts.h;
            P.ts[b].glin_method = ts.method;
        }
    }
    const double* v = P.v + (size_t)b * T.n;
    for (int i = threadIdx.x; i < T.n; i += blockDim.x) P.vI[IX(i, b, P.Bp)] = v[i];
    __syncthreads();
    if (P.p32) residual_to_interleaved(T, P, b, v);
    else
        for (int i = threadIdx.x; i < T.n; i += blockDim.x) P.rhslin[IX(i, b, P.Bp)] = rhs_lin_node(T, P, b, i);
    if (threadIdx.x == 0) P.st[b].done = DONE_RUNNING;
}

__device__ __forceinline__ double lte_norm(double v) {
#pragma clang fp contract(off)
    return 1e-6 + fabs(v) * 1e-3;
}

// calculateLTE (physics/circuitsim.h:177-246) of instance b at its Newton
// solution v: max over nodes, capacitor branches and MOSFET gate drives of
// the linear-predictor error times dt / (dt + prev_dt) over
// 1e-6 + 1e-3 |v|.  The maximum keeps std::max's order (a NaN term is
// skipped), so it does not depend on how the block splits the terms.
__device__ double block_lte(const DevTopo& T, const DevBatch& P, int b, const double* v, const TranState& ts,
                            double* red) {
#pragma clang fp contract(off)
    const int n = T.n;
    const double dt = ts.h, pdt = ts.prev_dt;
    if (!(dt > 0) || !(pdt > 0)) return 0.0;
    const double c_factor = dt / (dt + pdt);
    const double* p1 = P.arb_prev + (size_t)b * n;
    const double* p2 = P.arb_prev2 + (size_t)b * n;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v_pred = p1[i] + (p1[i] - p2[i]) * (dt / pdt);
        const double err = fabs(v[i] - v_pred) * c_factor;
        mx = ref_max(mx, err / lte_norm(v[i]));
    }
    for (int i = threadIdx.x; i < T.nC; i += blockDim.x) {
        const size_t ci = (size_t)b * T.nC + i;
        const int p = T.cap_n1[i], q = T.cap_n2[i];
        const double v1 = p > 0 ? v[p - 1] : 0.0;
        const double v2 = q > 0 ? v[q - 1] : 0.0;
        const double v_branch = v1 - v2;
        const double v_pred = P.hv0[ci] + (P.hv0[ci] - P.hv1[ci]) * (dt / pdt);
        const double err = fabs(v_branch - v_pred) * c_factor;
        mx = ref_max(mx, err / lte_norm(v_branch));
    }
    for (int i = threadIdx.x; i < T.nM; i += blockDim.x) {
        const int g = T.m_g[i], so = T.m_s[i];
        const double vgs = volt(v, n, g) - volt(v, n, so);
        const double vgs_p = volt(p1, n, g) - volt(p1, n, so);
        const double vgs_pp = volt(p2, n, g) - volt(p2, n, so);
        const double vgs_pred = vgs_p + (vgs_p - vgs_pp) * (dt / pdt);
        const double err = fabs(vgs - vgs_pred) * c_factor;
        mx = ref_max(mx, err / lte_norm(vgs));
    }
    return block_max(mx, red);
}

// Transient step control after an attempt's Newton solve (one block per
// instance whose attempt ended).  The restated stepTransient
// (solvers/circuitsim_solver.cpp:559-780) as its caller drives it
// (engine_impl.cpp:205-256):
//   - failed Newton solve: the fixed control ends the instance; the arbiter
//     halves h (at most 4 attempts a step);
//   - converged, arbiter control: analyzeStability (physics/circuitsim.h:
//     276-367): LTE > 1 or energy growth beyond E_prev + 1e-9 + 1.5 P_in h
//     + 1e-6 and 1.5 E_prev counts an instability: h halves, or at 3 or more
//     the method switches from trapezoidal to Gear-2 (same h), or, already
//     switched, the step is taken anyway; a stable step decrements the
//     counter and becomes the arbiter's new state (its rejectHistory never
//     alters the estimate, see the oracle);
//   - accepted: capacitor history shifted with the method's current
//     (:668-699), t += h, waveform ring entry every `decimation` steps, the
//     next step starts at h from the arbiter's voltages, or the instance is
//     finished (`steps` steps, or t >= t_stop).
// Runs inside every transient sweep over the active list, so each instance
// goes on to its next attempt without waiting for the others.
struct TranArgs {
    int control, steps, decimation, n_probe, wave_cap;
    double h0, t_stop;
    const int* probe;
    double *wave_t, *wave;
};

__global__ __launch_bounds__(256) void k_tran_control(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                      int nact_grid, TranArgs A) {
#pragma clang fp contract(off)
    __shared__ double red[8];
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    if ((int)blockIdx.x >= nact) return;
    const int b = act[blockIdx.x];
    const int control = A.control, steps = A.steps, decimation = A.decimation, n_probe = A.n_probe;
    const int wave_cap = A.wave_cap;
    const double h0 = A.h0, t_stop = A.t_stop;
    const int* __restrict__ probe = A.probe;
    double* __restrict__ wave_t = A.wave_t;
    double* __restrict__ wave = A.wave;
    const int n = T.n;
    InstState st = P.st[b];
    if (st.done != DONE_CONVERGED && st.done != DONE_FAILED) return;
    TranState ts = P.ts[b];
    double* v = P.v + (size_t)b * n;
    const bool arbiter = control == NR_TRAN_ARBITER;
    enum { ACCEPT, RETRY, FAIL };
    int decision;
    if (st.done == DONE_FAILED) {
        if (!arbiter) {
            decision = FAIL;
        } else {
            ts.rejected++;
            ts.h *= 0.5;
            ts.attempt++;
            decision = ts.attempt >= 4 ? FAIL : RETRY;
        }
    } else if (!arbiter) {
        decision = ACCEPT;
    } else {
        const double lte = block_lte(T, P, b, v, ts, red);
        const double energy = cap_energy(T, P, b, v);
        const double power_in = block_ordered_sum(T.nI, [&](int i) {
#pragma clang fp contract(off)
            double v_diff = 0.0;
            const int p = T.isrc_pos[i], q = T.isrc_neg[i];
            if (p > 0 && p <= n) v_diff += v[p - 1];
            if (q > 0 && q <= n) v_diff -= v[q - 1];
            return fabs(P.isrc[(size_t)b * T.nI + i] * v_diff);
        });
        const double tol = 1e-9 + power_in * ts.h * 1.5;
        const bool exploding = (energy > ts.prev_energy + tol + 1e-6) && (energy > ts.prev_energy * 1.5);
        int sug;
        if (lte > 1.0 || exploding) {
            ts.instab++;
            sug = ts.instab >= 3 ? 2 : 1;
        } else {
            ts.instab = max(0, ts.instab - 1);
            double* p1 = P.arb_prev + (size_t)b * n;
            double* p2 = P.arb_prev2 + (size_t)b * n;
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                p2[i] = p1[i];
                p1[i] = v[i];
            }
            ts.prev_energy = energy;
            ts.prev_dt = ts.h;
            sug = 0;
        }
        if (sug == 1) {
            ts.rejected++;
            ts.h *= 0.5;
            ts.attempt++;
            decision = ts.attempt >= 4 ? FAIL : RETRY;
        } else if (sug == 2 && ts.method == NR_INT_TRAP) {
            ts.rejected++;
            ts.method = NR_INT_GEAR2;
            ts.switched_at = ts.accepted;
            decision = RETRY;
        } else {
            decision = ACCEPT;
        }
    }
    if (decision == FAIL) ts.h = P.ts[b].h;   // the failed attempt's step
    bool finished = false;
    if (decision == ACCEPT) {
        for (int i = threadIdx.x; i < T.nC; i += blockDim.x) {
            const size_t ci = (size_t)b * T.nC + i;
            double vc = 0.0;
            const int p = T.cap_n1[i], q = T.cap_n2[i];
            if (p > 0 && p <= n) vc += v[p - 1];
            if (q > 0 && q <= n) vc -= v[q - 1];
            const double C = P.cap[ci], h = ts.h, v0 = P.hv0[ci], v1 = P.hv1[ci];
            double ic;
            if (ts.method == NR_INT_TRAP) ic = (2.0 * C / h) * (vc - v0) - P.hi0[ci];
            else if (ts.method == NR_INT_GEAR1) ic = (C / h) * (vc - v0);
            else if (v1 == 0.0 && ts.t > h) ic = (C / h) * (vc - v0);
            else ic = (C / h) * (1.5 * vc - 2.0 * v0 + 0.5 * v1);
            P.hv1[ci] = v0;
            P.hv0[ci] = vc;
            P.hi0[ci] = ic;
        }
        ts.t = ts.t + ts.h;
        ts.accepted++;
        if (++ts.dec >= decimation) {
            ts.dec = 0;
            const long long slot = ts.rec % wave_cap;
            if (threadIdx.x == 0) wave_t[(size_t)b * wave_cap + slot] = ts.t;
            double* w = wave + ((size_t)b * wave_cap + slot) * n_probe;
            for (int p = threadIdx.x; p < n_probe; p += blockDim.x) w[p] = volt(v, n, probe[p]);
            ts.rec++;
        }
        finished = arbiter ? !(ts.t < t_stop) : ts.accepted >= steps;
        if (!finished) {
            ts.h = h0;
            ts.attempt = 0;
        }
    }
    // the next attempt starts from the arbiter's voltages (the fixed control
    // continues from the accepted step)
    if (arbiter && (decision == RETRY || (decision == ACCEPT && !finished))) {
        __syncthreads();
        const double* p1 = P.arb_prev + (size_t)b * n;
        for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = p1[i];
    }
    if (threadIdx.x == 0) {
        st.step = ts.accepted;
        if (decision == FAIL) {
            st.done = DONE_FINISHED;
            if (st.status == NR_STATUS_CONVERGED) st.status = NR_STATUS_NOT_CONVERGED;
        } else if (decision == ACCEPT && finished) {
            st.done = DONE_FINISHED;
            st.status = NR_STATUS_CONVERGED;
        } else {
            st.done = DONE_ARMED;
            st.it = 0;
            st.singular = 0;
            st.prev_kcl = 0.0;
            st.status = NR_STATUS_NOT_CONVERGED;
        }
        P.st[b] = st;
        P.ts[b] = ts;
    }
}

// waveform ring entry 0: the start state at t = 0
__global__ __launch_bounds__(64) void k_wave_start(DevTopo T, DevBatch P, int n_probe, const int* __restrict__ probe,
                                                   int wave_cap, double* __restrict__ wave_t,
                                                   double* __restrict__ wave) {
    const int b = blockIdx.x;
    const double* v = P.v + (size_t)b * T.n;
    double* w = wave + (size_t)b * wave_cap * n_probe;
    for (int p = threadIdx.x; p < n_probe; p += blockDim.x) w[p] = volt(v, T.n, probe[p]);
    if (threadIdx.x == 0) {
        wave_t[(size_t)b * wave_cap] = 0.0;
        P.ts[b].rec = 1;
    }
}

// thread (item, lane): 4 items x 64 instances per block; blockIdx.y = group
#define LANE_SETUP(ITEMS_PER_BLOCK)                                       \
    const int item = blockIdx.x * (ITEMS_PER_BLOCK) + (threadIdx.x >> 6); \
    const int slot_ = blockIdx.y * 64 + (threadIdx.x & 63);               \
    if (slot_ >= nact) return;                                            \
    const int b = act[slot_];                                             \
    const int Bp = P.Bp;

__global__ __launch_bounds__(256) void k_eval(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid) {
#pragma clang fp contract(off)
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    LANE_SETUP(4)
    const int n = T.n;
    auto V = [&](int node) -> double { return (node > 0 && node <= n) ? P.vI[IX(node - 1, b, Bp)] : 0.0; };
    if (item < T.nD) {
        const int i = item;
        double vd = V(T.d_a[i]) - V(T.d_c[i]);
        double id, g;
        diode_eval(vd, P.IsI[IX(i, b, Bp)], T.d_n[i], diode_vt(T, P, i, b), id, g);
        P.dg[IX(i, b, Bp)] = g;
        P.dieq[IX(i, b, Bp)] = id - g * vd;
    } else if (item < T.nD + T.nM) {
        const int j = item - T.nD;
        const int d = T.m_d[j], g = T.m_g[j], s = T.m_s[j], bb = T.m_b[j];
        double vgs, vds, ids, gm, gds;
        mos_eval(V(g), V(d), V(s), T.m_w[j], T.m_l[j], P.kpI[IX(j, b, Bp)], P.vthI[IX(j, b, Bp)], T.m_lam[j],
                 T.m_pmos[j], vgs, vds, ids, gm, gds);
        const double gmb = 0.0;
        double vbs = (bb != 0) ? (V(bb) - (s != 0 ? V(s) : 0.0)) : 0.0;
        P.mgds[IX(j, b, Bp)] = gds;
        P.mgm[IX(j, b, Bp)] = gm;
        P.mieq[IX(j, b, Bp)] = ids - (gm * vgs + gmb * vbs + gds * vds);
    }
}

// Slots (and rhs rows) with more device terms than this are assembled by
// k_assemble_hub, which keeps k_assemble's register use low
constexpr int HUB_TERMS = 16;

// Sum of a term list in list order, starting from `acc`.  Short lists (the
// usual one or two device stamps of a slot) go term by term; long ones (a
// supply rail's diagonal holds one term per device) are loaded 16 at a time
// with indices clamped to the last term, so the loads of a batch are all in
// flight together.  load(t) returns the value of term t, add(acc, t, x) adds it.
template <typename A, typename Load, typename Add>
__device__ __forceinline__ A sum_terms(A acc, const int* __restrict__ terms, int k0, int k1, Load load, Add add) {
    if (k1 - k0 <= 8) {
        for (int k = k0; k < k1; ++k) {
            const int t = terms[k];
            acc = add(acc, t, load(t));
        }
        return acc;
    }
    int kb = k0;
    for (; k1 - kb > 4; kb += 16) {
        int tt[16];
        A x[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) tt[u] = terms[min(kb + u, k1 - 1)];
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = load(tt[u]);
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (kb + u < k1) acc = add(acc, tt[u], x[u]);
    }
    if (kb < k1) {
        int tt[4];
        A x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) tt[u] = terms[min(kb + u, k1 - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = load(tt[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (kb + u < k1) acc = add(acc, tt[u], x[u]);
    }
    return acc;
}

// A = linear part + device stamps (emission order) + gmin; rhs likewise.
// Each value goes to the front that owns it: the interleaved small-front pool
// or the instance-major large-front arrays.
__global__ __launch_bounds__(256) void k_assemble(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid,
                                                  const double* __restrict__ gmin_tab, int gmin_last) {
#pragma clang fp contract(off)
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    LANE_SETUP(4)
    if (item < T.nnz_a) {
        if (T.nl_ptr[item + 1] - T.nl_ptr[item] > HUB_TERMS) goto rhs;   // k_assemble_hub
        double a = P.glin[IX(item, b, Bp)];
        for (int k = T.nl_ptr[item]; k < T.nl_ptr[item + 1]; ++k) {
            const int t = T.nl_term[k];
            const int kind = term_kind(t);
            const double* src = kind == karam::SK_DIODE ? P.dg : kind == karam::SK_MOS_GDS ? P.mgds : P.mgm;
            const double g = src[IX(term_idx(t), b, Bp)];
            a = term_neg(t) ? a + (-g) : a + g;
        }
        if (T.slot_diag[item]) a = a + gmin_tab[min(P.st[b].step, gmin_last)];
        const int d = T.slot_dst[item];
        if (d & 1) P.Al[IX(d >> 1, b, Bp)] = a;
        else P.pool[IX(d >> 1, b, Bp)] = a;
    }
rhs:
    if (item < T.n) {
        if (T.rhs_nl_ptr[item + 1] - T.rhs_nl_ptr[item] > HUB_TERMS) return;   // k_assemble_hub
        double r = P.rhslin[IX(item, b, Bp)];
        for (int k = T.rhs_nl_ptr[item]; k < T.rhs_nl_ptr[item + 1]; ++k) {
            const int t = T.rhs_nl_term[k];
            const double x = (term_kind(t) == karam::TK_DIODE ? P.dieq : P.mieq)[IX(term_idx(t), b, Bp)];
            r = term_neg(t) ? r + (-x) : r + x;
        }
        const int d = T.node_dst[item];
        if (d & 1) P.rhsl[IX(d >> 1, b, Bp)] = r;
        else P.pool[IX(d >> 1, b, Bp)] = r;
    }
}

// FP32 path: the Jacobian entries of the generic device kernels
// (tensors/physics_tensors.cpp:19-104) in single precision.  Junction
// voltages are differenced in FP64 and rounded once, parameters are rounded to
// float; only the conductances are produced (the rhs is the FP64 KCL vector).
__device__ __forceinline__ float diode_g32(double vd64, double Is, double N, double Vt) {
#pragma clang fp contract(off)
    const float vd = (float)vd64, n_vt = (float)N * (float)Vt;
    const float arg = vd / n_vt;
    const float argc = (30.0f < arg) ? 30.0f : arg;
    const float gb = ((float)Is / n_vt) * expf(argc);
    return (1e-12f < gb) ? gb : 1e-12f;
}

__device__ __forceinline__ void mos_g32(double vgs64, double vds64, double W, double L, double Kp, double Vth,
                                        double lam, int pmos, float& gm, float& gds) {
#pragma clang fp contract(off)
    const float sign = pmos ? -1.0f : 1.0f;
    const float lvgs = sign * (float)vgs64, lvds = sign * (float)vds64, lvth = sign * (float)Vth;
    const float K = 0.5f * (float)Kp * ((float)W / (float)L);
    const float vov = lvgs - lvth;
    if (vov <= 0.0f) {
        gm = 0.0f; gds = 1e-12f;
    } else if (lvds < vov) {
        gm = 2.0f * K * lvds;
        gds = 2.0f * K * (vov - lvds);
    } else {
        gm = 2.0f * K * vov * (1.0f + (float)lam * lvds);
        gds = K * vov * vov * (float)lam;
    }
    if (gds < 1e-12f) gds = 1e-12f;
}

__global__ __launch_bounds__(256) void k_eval32(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid) {
#pragma clang fp contract(off)
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    LANE_SETUP(4)
    const int n = T.n;
    auto V = [&](int node) -> double { return (node > 0 && node <= n) ? P.vI[IX(node - 1, b, Bp)] : 0.0; };
    if (item < T.nD) {
        const int i = item;
        reinterpret_cast<float*>(P.dg)[IX(i, b, Bp)] =
            diode_g32(V(T.d_a[i]) - V(T.d_c[i]), P.IsI[IX(i, b, Bp)], T.d_n[i], diode_vt(T, P, i, b));
    } else if (item < T.nD + T.nM) {
        const int j = item - T.nD;
        const double vs = V(T.m_s[j]);
        float gm, gds;
        mos_g32(V(T.m_g[j]) - vs, V(T.m_d[j]) - vs, T.m_w[j], T.m_l[j], P.kpI[IX(j, b, Bp)], P.vthI[IX(j, b, Bp)],
                T.m_lam[j], T.m_pmos[j], gm, gds);
        reinterpret_cast<float*>(P.mgds)[IX(j, b, Bp)] = gds;
        reinterpret_cast<float*>(P.mgm)[IX(j, b, Bp)] = gm;
    }
}

// FP32 path: A in float (the linear part, summed in FP64, rounded once; then
// the FP32 device stamps and the gmin diagonal in emission order) and the
// Newton rhs b - A v = KCL vector - gmin v, formed in FP64 and rounded
__global__ __launch_bounds__(256) void k_assemble32(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid,
                                                    const double* __restrict__ gmin_tab, int gmin_last) {
#pragma clang fp contract(off)
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    LANE_SETUP(4)
    float* pool = reinterpret_cast<float*>(P.pool);
    const double gmin = gmin_tab[min(P.st[b].step, gmin_last)];
    if (item < T.nnz_a) {
        const float* dg = reinterpret_cast<const float*>(P.dg);
        const float* mgds = reinterpret_cast<const float*>(P.mgds);
        const float* mgm = reinterpret_cast<const float*>(P.mgm);
        if (T.nl_ptr[item + 1] - T.nl_ptr[item] > HUB_TERMS) goto rhs;   // k_assemble_hub
        float a = (float)P.glin[IX(item, b, Bp)];
        for (int k = T.nl_ptr[item]; k < T.nl_ptr[item + 1]; ++k) {
            const int t = T.nl_term[k];
            const int kind = term_kind(t);
            const float g = (kind == karam::SK_DIODE ? dg : kind == karam::SK_MOS_GDS ? mgds : mgm)[IX(term_idx(t), b, Bp)];
            a = term_neg(t) ? a + (-g) : a + g;
        }
        if (T.slot_diag[item]) a = a + (float)gmin;
        const int d = T.slot_dst[item];
        if (d & 1) reinterpret_cast<float*>(P.Al)[IX(d >> 1, b, Bp)] = a;
        else pool[IX(d >> 1, b, Bp)] = a;
    }
rhs:
    if (item < T.n) {
        const float r = (float)(P.rhslin[IX(item, b, Bp)] - gmin * P.vI[IX(item, b, Bp)]);
        const int d = T.node_dst[item];
        if (d & 1) reinterpret_cast<float*>(P.rhsl)[IX(d >> 1, b, Bp)] = r;
        else pool[IX(d >> 1, b, Bp)] = r;
    }
}

// The A entries and FP64 rhs rows with more than HUB_TERMS device terms
// (a supply rail), one thread per (item, instance): the same sums in the
// same order as k_assemble / k_assemble32, with the loads batched.
// Slots and rhs rows with many device terms (a supply rail's diagonal holds
// one per device on the rail), one block of 16 wavefronts per (item, 64
// instances): the wavefronts load a chunk of HUB_SUM_CHUNK terms at once into
// LDS, then the first adds the chunk in list order, lane by lane (each lane an
// instance), so every value is the sequential sum k_assemble would form.
constexpr int HUB_SUM_CHUNK = 256;
constexpr int HUB_SUM_THREADS = 1024;

template <typename A>
__device__ __forceinline__ A hub_sum(A acc, const int* __restrict__ terms, int k0, int k1, int lane, int w,
                                     bool on, const double* const* src, size_t b_off, int Bp, double* lds) {
    A(*buf)[64] = reinterpret_cast<A(*)[64]>(lds);
    constexpr int PER = HUB_SUM_CHUNK / (HUB_SUM_THREADS / 64);
    for (int base = k0; base < k1; base += HUB_SUM_CHUNK) {
        const int cnt = min(HUB_SUM_CHUNK, k1 - base);
        A x[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = w * PER + u;
            const int t = terms[base + min(q, cnt - 1)];
            const A g = reinterpret_cast<const A*>(src[term_kind(t) & 3])[IX(term_idx(t), 0, Bp) + b_off];
            x[u] = term_neg(t) ? -g : g;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) buf[w * PER + u][lane] = x[u];
        __syncthreads();
        if (w == 0 && on)
            for (int q = 0; q < cnt; ++q) acc = acc + buf[q][lane];
        __syncthreads();
    }
    return acc;
}

__global__ __launch_bounds__(HUB_SUM_THREADS) void k_assemble_hub(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                                 int nact_grid, const double* __restrict__ gmin_tab,
                                                                 int gmin_last) {
#pragma clang fp contract(off)
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    const int item = blockIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int slot_ = blockIdx.y * 64 + lane;
    const bool on = slot_ < nact;
    const int b = on ? act[slot_] : act[0];
    const int Bp = P.Bp;
    __shared__ double lds[HUB_SUM_CHUNK * 64];
    // term sources by kind: the slot terms' SK_DIODE / SK_MOS_GDS / SK_MOS_GM
    // values and the rhs terms' TK_DIODE / TK_MOS values (kind & 3 indexes)
    const double* slot_src[4];
    const double* row_src[4];
    slot_src[karam::SK_DIODE & 3] = P.dg;
    slot_src[karam::SK_MOS_GDS & 3] = P.mgds;
    slot_src[karam::SK_MOS_GM & 3] = P.mgm;
    row_src[karam::TK_DIODE & 3] = P.dieq;
    row_src[karam::TK_MOS & 3] = P.mieq;
    if (item < T.n_hub_slot) {
        const int sl = T.hub_slot[item];
        const int d = T.slot_dst[sl];
        if (P.p32) {
            float a = hub_sum<float>((float)P.glin[IX(sl, b, Bp)], T.nl_term, T.nl_ptr[sl], T.nl_ptr[sl + 1], lane,
                                     w, on, slot_src, (size_t)b, Bp, lds);
            if (w == 0 && on) {
                if (T.slot_diag[sl]) a = a + (float)gmin_tab[min(P.st[b].step, gmin_last)];
                if (d & 1) reinterpret_cast<float*>(P.Al)[IX(d >> 1, b, Bp)] = a;
                else reinterpret_cast<float*>(P.pool)[IX(d >> 1, b, Bp)] = a;
            }
        } else {
            double a = hub_sum<double>(P.glin[IX(sl, b, Bp)], T.nl_term, T.nl_ptr[sl], T.nl_ptr[sl + 1], lane, w,
                                       on, slot_src, (size_t)b, Bp, lds);
            if (w == 0 && on) {
                if (T.slot_diag[sl]) a = a + gmin_tab[min(P.st[b].step, gmin_last)];
                if (d & 1) P.Al[IX(d >> 1, b, Bp)] = a;
                else P.pool[IX(d >> 1, b, Bp)] = a;
            }
        }
    } else if (item < T.n_hub_slot + T.n_hub_row && !P.p32) {
        const int i = T.hub_row[item - T.n_hub_slot];
        const double r = hub_sum<double>(P.rhslin[IX(i, b, Bp)], T.rhs_nl_term, T.rhs_nl_ptr[i],
                                         T.rhs_nl_ptr[i + 1], lane, w, on, row_src, (size_t)b, Bp, lds);
        if (w == 0 && on) {
            const int d = T.node_dst[i];
            if (d & 1) P.rhsl[IX(d >> 1, b, Bp)] = r;
            else P.pool[IX(d >> 1, b, Bp)] = r;
        }
    }
}

// value-type helpers of the multifrontal kernels (V = double: the FP64 path;
// V = float: the FP32 path, same operation order; V = cdouble: AC analysis
// on complex admittances, same operation order with the complex operations
// below)
__device__ __forceinline__ double vfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float vfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double vabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float vabs(float x) { return __builtin_fabsf(x); }
// a pivot the elimination treats as zero
__device__ __forceinline__ bool vsmall(double p) { return !(vabs(p) >= 1e-25); }
__device__ __forceinline__ bool vsmall(float p) { return !(vabs(p) >= 1e-25f); }
__device__ __forceinline__ bool vsmall(cdouble p) { return !(cmag(p) >= 1e-25); }
// the factor buffers are sized for FP64; the FP32 path uses their first half,
// the complex path has buffers of twice the size
template <typename V>
__device__ __forceinline__ V* vp(double* p) { return reinterpret_cast<V*>(p); }
template <typename V>
__device__ __forceinline__ const long long* gp_offsets(const DevTopo& T) { return T.sn_gp_off; }
template <>
__device__ __forceinline__ const long long* gp_offsets<cdouble>(const DevTopo& T) { return T.sn_gp_off_c; }
// Front blocks (pivot rows, update blocks and their rhs parts) of the large
// side are stored front-major: front J's blocks for instances 0..B-1 lie
// side by side, so the teams of one launch touch one contiguous region.
__device__ __forceinline__ size_t fblk(long long off, int b, int size, int B) {
    return (size_t)off * (size_t)B + (size_t)b * (size_t)size;
}

// ---------------------------------------------------------------------------
// kernels: small fronts (lane = instance, front in VGPRs)
// ---------------------------------------------------------------------------
// Partial LU of one small supernode for one instance per lane.  R is the
// compile-time front capacity (>= r); all indexing into F is static.  The
// gather table holds, for the A entries and then for each child, one pool
// item per front position (the pool's zero row where nothing contributes),
// so every load is unconditional and the sum order matches the oracle's.
template <int R, typename V>
__device__ __forceinline__ int front_small(const DevTopo& T, const DevBatch& P, const SmallRec& e, int b, bool on) {
#pragma clang fp contract(off)
    constexpr int TB = R * R + R;
    const int Bp = P.Bp;
    const int c = e.c, m = e.m, r = c + m, f = e.f;
    const int nch = e.nch;
    const int* tb = T.tb + e.tb_off;
    const V* pool = vp<V>(P.pool);
    const int bb = on ? b : 0;
    V F[R * R];
    V fr[R];
#pragma unroll
    for (int p = 0; p < R * R; ++p) F[p] = V(0) + pool[IX(tb[p], bb, Bp)];
#pragma unroll
    for (int i = 0; i < R; ++i) fr[i] = V(0) + pool[IX(tb[R * R + i], bb, Bp)];
    for (int k = 0; k < nch; ++k) {
        const int* t2 = tb + (k + 1) * TB;
#pragma unroll
        for (int p = 0; p < R * R; ++p) F[p] += pool[IX(t2[p], bb, Bp)];
#pragma unroll
        for (int i = 0; i < R; ++i) fr[i] += pool[IX(t2[R * R + i], bb, Bp)];
    }
    int sing = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        if (k < c) {
            const V piv = F[k * R + k];
            if (vsmall(piv)) sing = 1;
#pragma unroll
            for (int i = k + 1; i < R; ++i) F[i * R + k] = F[i * R + k] / piv;
#pragma unroll
            for (int i = k + 1; i < R; ++i) {
                const V l = F[i * R + k];
#pragma unroll
                for (int j = k + 1; j < R; ++j) F[i * R + j] = vfma(-l, F[k * R + j], F[i * R + j]);
                fr[i] = vfma(-l, fr[k], fr[i]);
            }
        }
    }
    if (on) {
        const long long uo = e.u_off, co = e.cb_off, ro = e.rcb_off;
        const bool root = e.root != 0;
        V* cbl = vp<V>(P.CBL) + fblk(co, b, m * m, P.B);
        V* rcbl = vp<V>(P.RCBL) + fblk(ro, b, m, P.B);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (i < c) {
                vp<V>(P.ys)[IX(f + i, b, Bp)] = fr[i];
#pragma unroll
                for (int j = 0; j < R; ++j)
                    if (j < r) vp<V>(P.Us)[IX(uo + i * r + j, b, Bp)] = F[i * R + j];
            } else if (i < r) {
                if (root) rcbl[i - c] = fr[i];
                else vp<V>(P.pool)[IX(ro + (i - c), b, Bp)] = fr[i];
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    if (j >= c && j < r) {
                        if (root) cbl[(i - c) * m + (j - c)] = F[i * R + j];
                        else vp<V>(P.pool)[IX(co + (i - c) * m + (j - c), b, Bp)] = F[i * R + j];
                    }
                }
            }
        }
    }
    return sing;
}

template <int R, typename V>
__device__ __forceinline__ void back_small(const DevTopo& T, const DevBatch& P, const SmallRec& e, int b, bool on) {
#pragma clang fp contract(off)
    const int Bp = P.Bp;
    const int c = e.c, m = e.m, r = c + m, f = e.f;
    const long long uo = e.u_off;
    const int* rows = T.row_idx + e.row_off;
    const int bb = on ? b : 0;
    V t[R];
#pragma unroll
    for (int k = 0; k < R; ++k) t[k] = (k < c) ? vp<V>(P.ys)[IX(f + k, bb, Bp)] : V(0);
    for (int a = 0; a < m; ++a) {
        const V xa = vp<V>(P.xs)[IX(rows[a], bb, Bp)];
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (k < c) t[k] = vfma(-vp<V>(P.Us)[IX(uo + k * r + c + a, bb, Bp)], xa, t[k]);
    }
#pragma unroll
    for (int k = R - 1; k >= 0; --k) {
        if (k < c) {
            const V xk = t[k] / vp<V>(P.Us)[IX(uo + k * r + k, bb, Bp)];
            if (on) vp<V>(P.xs)[IX(f + k, b, Bp)] = xk;
#pragma unroll
            for (int i = 0; i < k; ++i) t[i] = vfma(-vp<V>(P.Us)[IX(uo + i * r + k, bb, Bp)], xk, t[i]);
        }
    }
}

// one wavefront per (small subtree, 64 instances); 4 wavefronts per block.
// Every front of the subtrees in one launch is padded to RMAX, so each
// kernel carries a single front shape (register budget and code size).
template <int RMAX, typename V>
__global__ __launch_bounds__(256) void k_factor_small(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid,
                                                      const int* __restrict__ sub_ptr,
                                                      const SmallRec* __restrict__ sub_rec, int n_sub) {
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    const int lane = threadIdx.x & 63;
    const long long task = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ngroups = (nact + 63) / 64;
    if (task >= (long long)n_sub * ngroups) return;
    const int sub = (int)(task % n_sub), g = (int)(task / n_sub);
    const int slot = g * 64 + lane;
    const bool on = slot < nact;
    const int b = on ? act[slot] : 0;
    int sing = 0;
    for (int q = sub_ptr[sub]; q < sub_ptr[sub + 1]; ++q) sing |= front_small<RMAX, V>(T, P, sub_rec[q], b, on);
    if (on && sing) P.st[b].singular = 1;
}

template <int RMAX, typename V>
__global__ __launch_bounds__(256) void k_backsolve_small(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                         int nact_grid, const int* __restrict__ sub_ptr,
                                                         const SmallRec* __restrict__ sub_rec, int n_sub) {
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    const int lane = threadIdx.x & 63;
    const long long task = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int ngroups = (nact + 63) / 64;
    if (task >= (long long)n_sub * ngroups) return;
    const int sub = (int)(task % n_sub), g = (int)(task / n_sub);
    const int slot = g * 64 + lane;
    const bool on = slot < nact;
    const int b = on ? act[slot] : 0;
    for (int q = sub_ptr[sub + 1] - 1; q >= sub_ptr[sub]; --q) back_small<RMAX, V>(T, P, sub_rec[q], b, on);
}

// ---------------------------------------------------------------------------
// kernels: large fronts (one team per (front, instance), front in LDS)
// ---------------------------------------------------------------------------
template <int TEAM>
__device__ __forceinline__ void team_sync() {
    if constexpr (TEAM <= 64) {
        // one wavefront: its LDS operations execute in order, so ordering
        // the compiler is enough (no wait on outstanding global loads)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Large fronts, one team per (front, instance).  The front's panel (the c
// pivot rows (c x r), the pivot columns below them (m x c) and the rhs) is
// assembled in LDS, or with GP in a scratch area in HBM (fronts whose panel
// exceeds the LDS a block holds), then eliminated in three phases that
// reproduce, element by element, the operations and their order of a
// right-looking elimination of the assembled front:
//   (B) one wavefront factors the c x c diagonal block and the pivot rows'
//       rhs, with wavefront-level barriers only;
//   (C) every lane takes a column of the pivot rows right of the block (U12)
//       or a row of the pivot columns below it (L21) and applies the c
//       updates to it alone: no barrier;
//   (D) the trailing m x m block (the front's update block) is formed in
//       4 x 2 register tiles and streamed to HBM, its entries gathered straight
//       from the children's update blocks (every A entry of a front lies in a
//       pivot row or column, so only children contribute).
// One front of one instance, by a team of TEAM lanes (lane = its index in
// the team) whose panel Pr (np = c r + m c + r values) lies in LDS or in
// HBM.  The panel is cleared first.  Every entry receives the same
// operations in the same order whatever TEAM is.
template <int TEAM, typename V>
__device__ __forceinline__ void factor_front(const DevTopo& T, const DevBatch& P, const FrontRec& fe, int b, int lane,
                                             V* Pr) {
#pragma clang fp contract(off)
    const int c = fe.c, m = fe.m, r = c + m, f = fe.f;
    const int np = c * r + m * c + r;                 // panel size
    // pivot rows, c x r (row stride r);
    // front-local (i, j) with i < c or j < c; rows c.. of the pivot columns
    // have stride c in the panel
    const int ls = c;
    V* Lc = Pr + c * r;                               // row i of L21 at Lc + i * ls
    V* fr = Pr + c * r + m * c;                       // rhs, r
    auto pidx = [&](int i, int j) { return i < c ? i * r + j : c * r + (i - c) * ls + j; };
    for (int e = lane; e < np; e += TEAM) Pr[e] = V(0);
    team_sync<TEAM>();
    // (A) assembly: A entries, the pivot rows' rhs, then each child in order
    const int Bp = P.Bp;
    const V* Al = vp<V>(P.Al);
    {
        const int a0 = fe.a_begin, a1 = fe.a_end;
        for (int e0 = a0 + lane; e0 < a1; e0 += 8 * TEAM) {
            V vals[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * TEAM;
                vals[u] = (e < a1) ? Al[IX(T.fa_slot[e], b, Bp)] : V(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * TEAM;
                if (e < a1) Pr[pidx(T.fa_li[e], T.fa_lj[e])] += vals[u];
            }
        }
    }
    const V* rl = vp<V>(P.rhsl);
    for (int k = lane; k < c; k += TEAM) fr[k] += rl[IX(T.perm[f + k], b, Bp)];
    team_sync<TEAM>();
    const int ch0 = fe.ch_begin, ch1 = fe.ch_end;
    for (int ch = ch0; ch < ch1; ++ch) {
        const ChildRec cr = T.child_rec[ch];
        const int mc = cr.nrows, pc = cr.pc;
        const int* rel = T.relmap + cr.row_off;
        const V* cb = vp<V>(P.CBL) + fblk(cr.cb_off, b, mc * mc, P.B);
        const V* rcb = vp<V>(P.RCBL) + fblk(cr.rcb_off, b, mc, P.B);
        // rows a < pc (all columns) and columns q < pc of the remaining rows:
        // the child's entries that land in the panel
        const int n1 = pc * mc, nn = n1 + (mc - pc) * pc;
        for (int e0 = lane; e0 < nn; e0 += 8 * TEAM) {
            V vals[8];
            int pos[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * TEAM;
                int a = 0, q = 0;
                if (e < n1) { a = e / mc; q = e - a * mc; }
                else if (e < nn) { const int e2 = e - n1; a = pc + e2 / pc; q = e2 - (a - pc) * pc; }
                vals[u] = (e < nn) ? cb[a * mc + q] : V(0);
                pos[u] = (e < nn) ? pidx(rel[a], rel[q]) : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + u * TEAM < nn) Pr[pos[u]] += vals[u];
        }
        for (int a = lane; a < mc; a += TEAM) fr[rel[a]] += rcb[a];
        team_sync<TEAM>();
    }
    // (B) diagonal block and pivot-row rhs: at step k every lane owns a
    // column j > k (SW consecutive columns per row stripe) and a stripe of the
    // rows i > k, so the multipliers are broadcast reads and the row loop
    // runs four independent updates at a time
    constexpr int SW = TEAM < 64 ? TEAM : 64;
    constexpr int NS = TEAM / SW;
    const int jl = lane % SW, st = lane / SW;
    int sing = 0;
    for (int k = 0; k < c; ++k) {
        const V piv = Pr[k * r + k];
        if (vsmall(piv)) sing = 1;
        for (int i = k + 1 + lane; i < c; i += TEAM) Pr[i * r + k] = Pr[i * r + k] / piv;
        team_sync<TEAM>();
        for (int j = k + 1 + jl; j < c; j += SW) {
            const V u = Pr[k * r + j];
            int i = k + 1 + st;
            for (; i + 3 * NS < c; i += 4 * NS) {
                const V l0 = Pr[i * r + k], l1 = Pr[(i + NS) * r + k];
                const V l2 = Pr[(i + 2 * NS) * r + k], l3 = Pr[(i + 3 * NS) * r + k];
                const V a0 = Pr[i * r + j], a1 = Pr[(i + NS) * r + j];
                const V a2 = Pr[(i + 2 * NS) * r + j], a3 = Pr[(i + 3 * NS) * r + j];
                Pr[i * r + j] = vfma(-l0, u, a0);
                Pr[(i + NS) * r + j] = vfma(-l1, u, a1);
                Pr[(i + 2 * NS) * r + j] = vfma(-l2, u, a2);
                Pr[(i + 3 * NS) * r + j] = vfma(-l3, u, a3);
            }
            for (; i < c; i += NS) Pr[i * r + j] = vfma(-Pr[i * r + k], u, Pr[i * r + j]);
        }
        for (int i = k + 1 + lane; i < c; i += TEAM) fr[i] = vfma(-Pr[i * r + k], fr[k], fr[i]);
        team_sync<TEAM>();
    }
    // (C) U12 columns and L21 rows, each by one lane
    for (int t = lane; t < 2 * m; t += TEAM) {
        if (t < m) {
            const int j = c + t;
            for (int k = 1; k < c; ++k) {
                V x = Pr[k * r + j];
                for (int p = 0; p < k; ++p) x = vfma(-Pr[k * r + p], Pr[p * r + j], x);
                Pr[k * r + j] = x;
            }
        } else {
            V* Li = Lc + (t - m) * ls;
            for (int k = 0; k < c; ++k) {
                V x = Li[k];
                for (int p = 0; p < k; ++p) x = vfma(-Li[p], Pr[p * r + k], x);
                Li[k] = x / Pr[k * r + k];
            }
        }
    }
    team_sync<TEAM>();
    // (D) trailing block in TI x TJ register tiles, streamed to the update block
    if (m > 0) {
        constexpr int TI = 4, TJ = 2;
        V* cbJ = vp<V>(P.CBL) + fblk(fe.cb_off, b, m * m, P.B);
        const int mti = (m + TI - 1) / TI, mtj = (m + TJ - 1) / TJ;
        for (int tt = lane; tt < mti * mtj; tt += TEAM) {
            const int i0 = (tt / mtj) * TI, j0 = (tt % mtj) * TJ;
            V acc[TI][TJ];
#pragma unroll
            for (int a = 0; a < TI; ++a)
#pragma unroll
                for (int q = 0; q < TJ; ++q) acc[a][q] = V(0);
            for (int ch = ch0; ch < ch1; ++ch) {
                const ChildRec cr = T.child_rec[ch];
                const int mc = cr.nrows;
                const int* inv = T.inv + cr.inv_off + c;
                const V* cb = vp<V>(P.CBL) + fblk(cr.cb_off, b, mc * mc, P.B);
                int ra[TI], qa[TJ];
#pragma unroll
                for (int a = 0; a < TI; ++a) ra[a] = (i0 + a < m) ? inv[i0 + a] : -1;
#pragma unroll
                for (int q = 0; q < TJ; ++q) qa[q] = (j0 + q < m) ? inv[j0 + q] : -1;
                V x[TI][TJ];
#pragma unroll
                for (int a = 0; a < TI; ++a)
#pragma unroll
                    for (int q = 0; q < TJ; ++q)
                        x[a][q] = (ra[a] >= 0 && qa[q] >= 0) ? cb[ra[a] * mc + qa[q]] : V(0);
#pragma unroll
                for (int a = 0; a < TI; ++a)
#pragma unroll
                    for (int q = 0; q < TJ; ++q)
                        if (ra[a] >= 0 && qa[q] >= 0) acc[a][q] += x[a][q];
            }
            for (int k = 0; k < c; ++k) {
                V l[TI], u[TJ];
#pragma unroll
                for (int a = 0; a < TI; ++a) l[a] = (i0 + a < m) ? Lc[(i0 + a) * ls + k] : V(0);
#pragma unroll
                for (int q = 0; q < TJ; ++q) u[q] = (j0 + q < m) ? Pr[k * r + c + j0 + q] : V(0);
#pragma unroll
                for (int a = 0; a < TI; ++a)
#pragma unroll
                    for (int q = 0; q < TJ; ++q) acc[a][q] = vfma(-l[a], u[q], acc[a][q]);
            }
#pragma unroll
            for (int a = 0; a < TI; ++a)
#pragma unroll
                for (int q = 0; q < TJ; ++q)
                    if (i0 + a < m && j0 + q < m) cbJ[(i0 + a) * m + j0 + q] = acc[a][q];
        }
        V* rcb = vp<V>(P.RCBL) + fblk(fe.rcb_off, b, m, P.B);
        for (int i = lane; i < m; i += TEAM) {
            V acc = fr[c + i];
            for (int k = 0; k < c; ++k) acc = vfma(-Lc[i * ls + k], fr[k], acc);
            rcb[i] = acc;
        }
    }
    V* U = vp<V>(P.UL) + fblk(fe.u_off, b, c * r, P.B);
    for (int e = lane; e < c * r; e += TEAM) U[e] = Pr[e];
    for (int k = lane; k < c; k += TEAM) vp<V>(P.ys)[IX(f + k, b, Bp)] = fr[k];
    if (sing && lane == 0) P.st[b].singular = 1;
}

// One launch per level and team class: a team per (front, instance)
template <int TEAM, int NT, bool GP, typename V>
__global__ __launch_bounds__(NT) void k_factor_large(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid,
                                                     const FrontRec* __restrict__ fr_list, int n_sn, int lds_team) {
    // a team wider than a wavefront synchronises the whole block, so it must
    // be alone in it (early exits and per-front barrier counts then agree)
    static_assert(TEAM <= 64 || NT == TEAM, "teams above 64 lanes need one team per block");
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* smem = reinterpret_cast<V*>(smem_raw);
    constexpr int TPB = NT / TEAM;
    const int tib = threadIdx.x / TEAM;
    const int lane = threadIdx.x % TEAM;
    // tasks are front-major: the teams sharing a wavefront factor the same
    // front for consecutive instances, so their control flow is identical
    constexpr int TPW = TEAM < 64 ? 64 / TEAM : 1;
    const int nact_pad = (nact + TPW - 1) / TPW * TPW;
    const long long task = (long long)blockIdx.x * TPB + tib;
    if (task >= (long long)n_sn * nact_pad) return;
    const FrontRec fe = fr_list[task / nact_pad];
    const int slot = (int)(task % nact_pad);
    if (slot >= nact) return;
    const int b = act[slot];
    const int np = fe.c * (fe.c + fe.m) + fe.m * fe.c + fe.c + fe.m;
    V* Pr = GP ? vp<V>(P.gpanel) + fblk(gp_offsets<V>(T)[fe.J], b, np, P.B)
               : smem + (size_t)tib * (size_t)lds_team;
    factor_front<TEAM, V>(T, P, fe, b, lane, Pr);
}

// Large fronts with the whole front in LDS, one team per (front, instance):
// the r x r front and its rhs are assembled there (the A entries, then each
// child's update block in child order, read once in its own row-major order
// so its cache lines are fetched once), the c pivots are eliminated
// right-looking in LDS, and the pivot rows, the update block and the rhs are
// streamed out in the layouts k_factor_large writes.  Every entry receives
// the same operations in the same order as in k_factor_large (A entry, the
// children in order, then one fma per pivot in pivot order), so results are
// identical; only the traffic and the synchronisation differ.  For fronts
// whose r (r + 1) values fit a team's LDS share.
template <int TEAM, int NT, typename V>
__global__ __launch_bounds__(NT) void k_factor_full(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid,
                                                    const FrontRec* __restrict__ fr_list, int n_sn, int lds_team) {
#pragma clang fp contract(off)
    static_assert(TEAM <= 64 || NT == TEAM, "teams above 64 lanes need one team per block");
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* smem = reinterpret_cast<V*>(smem_raw);
    constexpr int TPB = NT / TEAM;
    const int tib = threadIdx.x / TEAM;
    const int lane = threadIdx.x % TEAM;
    constexpr int TPW = TEAM < 64 ? 64 / TEAM : 1;
    const int nact_pad = (nact + TPW - 1) / TPW * TPW;
    const long long task = (long long)blockIdx.x * TPB + tib;
    if (task >= (long long)n_sn * nact_pad) return;
    const FrontRec fe = fr_list[task / nact_pad];
    const int slot = (int)(task % nact_pad);
    if (slot >= nact) return;
    const int b = act[slot];
    const int c = fe.c, m = fe.m, r = c + m, f = fe.f;
    V* F = smem + (size_t)tib * (size_t)lds_team;     // r x r, row-major
    V* fr = F + r * r;                                // rhs, r
    for (int e = lane; e < r * r + r; e += TEAM) F[e] = V(0);
    team_sync<TEAM>();
    const int Bp = P.Bp;
    const V* Al = vp<V>(P.Al);
    {
        const int a0 = fe.a_begin, a1 = fe.a_end;
        for (int e0 = a0 + lane; e0 < a1; e0 += 8 * TEAM) {
            V vals[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * TEAM;
                vals[u] = (e < a1) ? Al[IX(T.fa_slot[e], b, Bp)] : V(0);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * TEAM;
                if (e < a1) F[T.fa_li[e] * r + T.fa_lj[e]] += vals[u];
            }
        }
    }
    const V* rl = vp<V>(P.rhsl);
    for (int k = lane; k < c; k += TEAM) fr[k] += rl[IX(T.perm[f + k], b, Bp)];
    team_sync<TEAM>();
    const int ch0 = fe.ch_begin, ch1 = fe.ch_end;
    for (int ch = ch0; ch < ch1; ++ch) {
        const int C = T.children[ch];
        const int mc = T.sn_nrows[C];
        const int* rel = T.relmap + T.sn_row_off[C];
        const V* cb = vp<V>(P.CBL) + fblk(T.sn_cb_off[C], b, mc * mc, P.B);
        const V* rcb = vp<V>(P.RCBL) + fblk(T.sn_rcb_off[C], b, mc, P.B);
        const int nn = mc * mc;
        for (int e0 = lane; e0 < nn; e0 += 8 * TEAM) {
            V vals[8];
            int pos[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + u * TEAM;
                const int a = e < nn ? e / mc : 0, q = e < nn ? e - a * mc : 0;
                vals[u] = (e < nn) ? cb[e] : V(0);
                pos[u] = (e < nn) ? rel[a] * r + rel[q] : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + u * TEAM < nn) F[pos[u]] += vals[u];
        }
        for (int a = lane; a < mc; a += TEAM) fr[rel[a]] += rcb[a];
        team_sync<TEAM>();
    }
    int sing = 0;
    for (int k = 0; k < c; ++k) {
        const V piv = F[k * r + k];
        if (vsmall(piv)) sing = 1;
        for (int i = k + 1 + lane; i < r; i += TEAM) F[i * r + k] = F[i * r + k] / piv;
        team_sync<TEAM>();
        const int w = r - k - 1;
        for (int e = lane; e < w * w; e += TEAM) {
            const int i = k + 1 + e / w, j = k + 1 + e % w;
            F[i * r + j] = vfma(-F[i * r + k], F[k * r + j], F[i * r + j]);
        }
        for (int i = k + 1 + lane; i < r; i += TEAM) fr[i] = vfma(-F[i * r + k], fr[k], fr[i]);
        team_sync<TEAM>();
    }
    if (m > 0) {
        V* cbJ = vp<V>(P.CBL) + fblk(fe.cb_off, b, m * m, P.B);
        for (int e = lane; e < m * m; e += TEAM) {
            const int i = e / m, j = e - i * m;
            cbJ[e] = F[(c + i) * r + c + j];
        }
        V* rcbJ = vp<V>(P.RCBL) + fblk(fe.rcb_off, b, m, P.B);
        for (int i = lane; i < m; i += TEAM) rcbJ[i] = fr[c + i];
    }
    V* U = vp<V>(P.UL) + fblk(fe.u_off, b, c * r, P.B);
    for (int e = lane; e < c * r; e += TEAM) U[e] = F[e];
    for (int k = lane; k < c; k += TEAM) vp<V>(P.ys)[IX(f + k, b, Bp)] = fr[k];
    if (sing && lane == 0) P.st[b].singular = 1;
}

// Backward substitution of one front of one instance by a team of TEAM
// lanes, with the front's r solution values staged in xl (LDS)
template <int TEAM, typename V>
__device__ __forceinline__ void backsolve_front(const DevTopo& T, const DevBatch& P, const FrontRec& fe, int b,
                                                int lane, V* xl) {
#pragma clang fp contract(off)
    const int Bp = P.Bp;
    const int c = fe.c, m = fe.m, r = c + m, f = fe.f;
    const V* U = vp<V>(P.UL) + fblk(fe.u_off, b, c * r, P.B);
    const int* rows = T.row_idx + fe.row_off;
    V* xs = vp<V>(P.xs);
    const V* ys = vp<V>(P.ys);
    for (int a = lane; a < m; a += TEAM) xl[c + a] = xs[IX(rows[a], b, Bp)];
    team_sync<TEAM>();
    for (int k = lane; k < c; k += TEAM) {
        V s = ys[IX(f + k, b, Bp)];
        const V* Uk = U + (size_t)k * r + c;
        for (int a = 0; a < m; ++a) s = vfma(-Uk[a], xl[c + a], s);
        xl[k] = s;
    }
    team_sync<TEAM>();
    for (int k = c - 1; k >= 0; --k) {
        const V xk = xl[k] / U[(size_t)k * r + k];
        for (int i = lane; i < k; i += TEAM) xl[i] = vfma(-U[(size_t)i * r + k], xk, xl[i]);
        if (lane == 0) xs[IX(f + k, b, Bp)] = xk;
        team_sync<TEAM>();
    }
}

template <int TEAM, int NT, typename V>
__global__ __launch_bounds__(NT) void k_backsolve_large(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                        int nact_grid, const FrontRec* __restrict__ fr_list, int n_sn,
                                                        int rmax) {
    // a team wider than a wavefront synchronises the whole block, so it must
    // be alone in it (early exits and per-front barrier counts then agree)
    static_assert(TEAM <= 64 || NT == TEAM, "teams above 64 lanes need one team per block");
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* smem = reinterpret_cast<V*>(smem_raw);
    constexpr int TPB = NT / TEAM;
    const int tib = threadIdx.x / TEAM;
    const int lane = threadIdx.x % TEAM;
    // tasks are front-major: the teams sharing a wavefront factor the same
    // front for consecutive instances, so their control flow is identical
    constexpr int TPW = TEAM < 64 ? 64 / TEAM : 1;
    const int nact_pad = (nact + TPW - 1) / TPW * TPW;
    const long long task = (long long)blockIdx.x * TPB + tib;
    if (task >= (long long)n_sn * nact_pad) return;
    const FrontRec fe = fr_list[task / nact_pad];
    const int slot = (int)(task % nact_pad);
    if (slot >= nact) return;
    backsolve_front<TEAM, V>(T, P, fe, act[slot], lane, smem + (size_t)tib * (size_t)rmax);
}

// ---------------------------------------------------------------------------
// kernels: the whole large-front phase in one launch
// ---------------------------------------------------------------------------
// Every large front of an instance, level by level, by one block: each of
// its four wavefronts is a 64-lane team that takes the level's fronts in turn,
// and a block barrier separates the levels, so a front's children are
// complete and visible to the block before it starts (workgroup scope: no
// waits across blocks).  Blocks loop over the active instances.  A front
// whose panel exceeds the team's LDS share lds_team keeps its panel in HBM
// (sn_pgp_off).  The results are those of the per-level launches: the
// arithmetic of a front does not depend on the team.
template <typename V>
__global__ __launch_bounds__(256) void k_factor_persist(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                        int nact_grid, const FrontRec* __restrict__ list,
                                                        const int* __restrict__ lev_ptr, int n_lev, int lds_team) {
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    V* Pl = reinterpret_cast<V*>(smem_raw) + (size_t)wave * (size_t)lds_team;
    for (int slot = blockIdx.x; slot < nact; slot += gridDim.x) {
        const int b = act[slot];
        for (int l = 0; l < n_lev; ++l) {
            const int p1 = lev_ptr[l + 1];
            for (int p = lev_ptr[l] + wave; p < p1; p += 4) {
                const FrontRec fe = list[p];
                const int np = fe.c * (fe.c + fe.m) + fe.m * fe.c + fe.c + fe.m;
                if (np <= lds_team)
                    factor_front<64, V>(T, P, fe, b, lane, Pl);
                else
                    factor_front<64, V>(T, P, fe, b, lane,
                                        vp<V>(P.pgpanel) + fblk(T.sn_pgp_off[fe.J], b, np, P.B));
            }
            __syncthreads();
        }
    }
}

// The backward substitution of every large front of an instance, levels in
// reverse, by one block (as k_factor_persist); rmax = the largest front order
template <typename V>
__global__ __launch_bounds__(256) void k_backsolve_persist(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                           int nact_grid, const FrontRec* __restrict__ list,
                                                           const int* __restrict__ lev_ptr, int n_lev, int rmax) {
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    V* xl = reinterpret_cast<V*>(smem_raw) + (size_t)wave * (size_t)rmax;
    for (int slot = blockIdx.x; slot < nact; slot += gridDim.x) {
        const int b = act[slot];
        for (int l = n_lev - 1; l >= 0; --l) {
            const int p1 = lev_ptr[l + 1];
            for (int p = lev_ptr[l] + wave; p < p1; p += 4) backsolve_front<64, V>(T, P, list[p], b, lane, xl);
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// kernels: layout bridges
// ---------------------------------------------------------------------------
// xs (interleaved, permuted rows) -> xm (instance-major, permuted rows, FP64)
template <typename V>
__global__ __launch_bounds__(256) void k_x_to_xm(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid) {
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    __shared__ double tile[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64, g = blockIdx.y;
    const int slot = g * 64 + lane;
    const int n = T.n, Bp = P.Bp;
    if (slot < nact) {
        const int b = act[slot];
        const V* xs = vp<V>(P.xs);
        for (int rr = w; rr < 64; rr += 4)
            if (row0 + rr < n) tile[rr][lane] = (double)xs[IX(row0 + rr, b, Bp)];
    }
    __syncthreads();
    for (int ii = w; ii < 64; ii += 4) {
        const int s2 = g * 64 + ii;
        if (s2 >= nact) break;
        const int b = act[s2];
        if (row0 + lane < n) P.xm[(size_t)b * n + row0 + lane] = tile[lane][ii];
    }
}

// v (instance-major) -> vI (interleaved); on the FP32 path also the KCL
// vector res -> rhslin
__global__ __launch_bounds__(256) void k_v_to_vI(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid) {
    // eager launches pass the active count; graph replays pass minus the
    // grid's count and read the live one (at most that) from the device
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    __shared__ double tile[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64, g = blockIdx.y;
    const int n = T.n, Bp = P.Bp;
    const int slot = g * 64 + lane;
    for (int pass = 0; pass < (P.p32 ? 2 : 1); ++pass) {
        const double* src = pass ? P.res : P.v;
        double* dst = pass ? P.rhslin : P.vI;
        if (pass) __syncthreads();
        for (int ii = w; ii < 64; ii += 4) {
            const int s2 = g * 64 + ii;
            if (s2 >= nact) break;
            const int b = act[s2];
            if (row0 + lane < n) tile[lane][ii] = src[(size_t)b * n + row0 + lane];
        }
        __syncthreads();
        if (slot < nact) {
            const int b = act[slot];
            for (int rr = w; rr < 64; rr += 4)
                if (row0 + rr < n) dst[IX(row0 + rr, b, Bp)] = tile[rr][lane];
        }
    }
}

// ---------------------------------------------------------------------------
// kernels: the pivoting fallback of the linear solve
// ---------------------------------------------------------------------------
// The multifrontal LU does not pivot.  Where a circuit's Jacobian needs row
// exchanges for stability (a long inverter chain: each stage's gm exceeds
// its output conductance by the stage gain, and eliminating the chain in
// order multiplies the supply column by that gain at every stage), its
// element growth leaves a solution whose residual is orders of magnitude
// above the right-hand side.  NR_LU_AUTO measures every solve's relative
// residual max|b - A x| / max|b| in FP64 and hands the instances above the
// bound to the dense LU with partial pivoting that the reference runs
// (solveLU_Pivoted, math/linalg.h:742-854, reached from solvers/
// circuitsim_solver.cpp:291-295); NR_LU_DENSE uses that LU for every solve.
constexpr int LINRES_ROWS = 16;          // rows per wavefront of k_linres
constexpr double LU_ETA_FP64 = 1e-8;     // bound of the FP64 path's ratio
// The FP32 path's float system: its Newton step is an inexact Newton step
// (the FP64 residual corrects it), which needs the float solve to cut the
// linear residual, not to match it to rounding.  Accurate float solves of a
// 20,000-node logic circuit near its root sit at 1e-3..2e-2 (max|b| is small
// there); solves the missing pivoting spoils sit at 1e16 and above.
constexpr double LU_ETA_FP32 = 0.1;
constexpr int DENSE_THREADS = 1024;
// blocked dense LU: widest panel (pivot steps per pass over the trailing
// matrix) and the dynamic LDS of k_dense (the rhs, then the panel; static
// LDS stays below 2 KB of the 160 KB a workgroup may hold)
constexpr int DENSE_NB_MAX = 16;
constexpr int DENSE_ROWS = 4;      // trailing rows per lane in flight
constexpr size_t DENSE_LDS = 160 * 1024 - 2048;

// NR_LU_AUTO, after the back substitution: per instance, max_i |b_i - (A x)_i|
// and max_i |b_i| from the assembled values (the factorisation leaves the A
// entries and the rhs in place) and the solution xs, every row's residual
// accumulated in CSR order in FP64.  Lane = instance; each wavefront covers
// LINRES_ROWS rows, the block's four are combined in LDS and merged into the
// instance's maxima with one atomic each.
template <typename V>
__global__ __launch_bounds__(256) void k_linres(DevTopo T, DevBatch P, const int* __restrict__ act, int nact_grid) {
#pragma clang fp contract(off)
    const int nact = nact_grid > 0 ? nact_grid : *P.nact;
    __shared__ double smr[4][64], smb[4][64];
    __shared__ int sbad[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int slot = blockIdx.y * 64 + lane;
    const bool on = slot < nact;
    const int b = on ? act[slot] : 0, Bp = P.Bp;
    double mr = 0.0, mb = 0.0;
    int bad = 0;
    if (on) {
        const V* pool = vp<V>(P.pool);
        const V* Al = vp<V>(P.Al);
        const V* rl = vp<V>(P.rhsl);
        const V* xs = vp<V>(P.xs);
        const int i0 = (blockIdx.x * 4 + w) * LINRES_ROWS, i1 = min(i0 + LINRES_ROWS, T.n);
        for (int i = i0; i < i1; ++i) {
            const int dn = T.node_dst[i];
            const double bi = (double)((dn & 1) ? rl[IX(dn >> 1, b, Bp)] : pool[IX(dn >> 1, b, Bp)]);
            double r = bi;
            for (int q = T.row_ptr[i]; q < T.row_ptr[i + 1]; ++q) {
                const int d = T.slot_dst[q];
                const double a = (double)((d & 1) ? Al[IX(d >> 1, b, Bp)] : pool[IX(d >> 1, b, Bp)]);
                r = r - a * (double)xs[IX(T.iperm[T.col_idx[q]], b, Bp)];
            }
            const double xi = (double)xs[IX(T.iperm[i], b, Bp)];
            if (!isfinite(r) || !isfinite(xi) || !isfinite(bi)) bad = 1;
            if (fabs(r) > mr) mr = fabs(r);
            if (fabs(bi) > mb) mb = fabs(bi);
        }
    }
    smr[w][lane] = mr;
    smb[w][lane] = mb;
    sbad[w][lane] = bad;
    __syncthreads();
    if (w == 0 && on && blockIdx.x * 4 * LINRES_ROWS < T.n) {
        for (int u = 1; u < 4; ++u) {
            if (smr[u][lane] > mr) mr = smr[u][lane];
            if (smb[u][lane] > mb) mb = smb[u][lane];
            bad |= sbad[u][lane];
        }
        // non-negative doubles order as their bit patterns
        atomicMax(&P.dn_acc[2 * (size_t)b], (unsigned long long)__double_as_longlong(mr));
        atomicMax(&P.dn_acc[2 * (size_t)b + 1], (unsigned long long)__double_as_longlong(mb));
        if (bad) atomicOr(&P.dn_bad[b], 1);
    }
}

// NR_LU_AUTO: list the active instances whose solve the dense LU redoes (a
// value that is not finite, a residual above bound * max|b|, or a pivot the
// multifrontal LU found below 1e-25), clear the maxima for the next sweep and
// rewind the queue.  With P.promote such instances are marked singular
// instead (k_update sends them to the FP64 path) and the list stays empty.
// One block.
__global__ __launch_bounds__(1024) void k_lu_flag(DevBatch P, const int* __restrict__ act, double bound) {
    __shared__ int cnt;
    const int nact = *P.nact;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (int s = threadIdx.x; s < nact; s += blockDim.x) {
        const int b = act[s];
        const double mr = __longlong_as_double((long long)P.dn_acc[2 * (size_t)b]);
        const double mb = __longlong_as_double((long long)P.dn_acc[2 * (size_t)b + 1]);
        const int bad = P.dn_bad[b];
        P.dn_acc[2 * (size_t)b] = 0ull;
        P.dn_acc[2 * (size_t)b + 1] = 0ull;
        P.dn_bad[b] = 0;
        if (bad || !(mr <= bound * mb) || P.st[b].singular) {
            if (P.promote) P.st[b].singular = 1;
            else P.dn_list[atomicAdd(&cnt, 1)] = b;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        P.dn_ctl[0] = cnt;
        P.dn_ctl[1] = 0;
    }
}

// NR_LU_DENSE: every active instance goes to the dense LU.  One block.
__global__ __launch_bounds__(1024) void k_dense_all(DevBatch P, const int* __restrict__ act) {
    const int nact = *P.nact;
    for (int s = threadIdx.x; s < nact; s += blockDim.x) P.dn_list[s] = act[s];
    if (threadIdx.x == 0) {
        P.dn_ctl[0] = nact;
        P.dn_ctl[1] = 0;
    }
}

// larger |value|, then the lower row: the pivot rule of solveLU_Pivoted
__device__ __forceinline__ void piv_better(double& v, int& k, double v2, int k2) {
    if (v2 > v || (v2 == v && k2 < k)) { v = v2; k = k2; }
}

// solveLU_Pivoted (math/linalg.h:742-854) on the listed instances, one
// workgroup per instance taken from the queue, `dn_slots` workgroups.  The
// system is the assembled A (as double) and rhs in the original node order,
// dense and row-major in the workgroup's workspace.  Every element receives
// the reference's operations in the reference's order: for each pivot step
// the pivot is the largest |a| in the column (NaN never, ties to the lower
// row; below 1e-25 the system is singular), whole rows and rhs entries are
// swapped, each lower row's multiplier is a / pivot and, when not zero, its
// trailing entries take a - l * u; the forward substitution applies the
// columns to each rhs entry in increasing order, and the backward
// substitution subtracts each row's products in increasing column order
// before dividing by the diagonal.  So the solution is bit-identical to the
// sequential routine.
// The elimination is blocked without changing any element's operations: a
// panel of w columns (w <= DENSE_NB_MAX, as many as the LDS holds for the
// remaining rows) is factored in LDS step by step (pivot search, the swap
// of the panel part of the two rows and of the rhs, multipliers, the
// panel's own updates); then, one thread per column outside the panel, the
// panel's row swaps are applied in step order to the rest of each row, each
// pivot row's entries right of the panel take the earlier pivot rows'
// updates in step order, and every trailing entry takes the panel's w
// updates in step order, one pass over the trailing matrix per panel
// instead of one per step.  An entry's value depends only on the updates it
// receives and their order, and a row's multipliers move with it, so the
// results are those of the unblocked routine bit for bit.
// The solution is written to xm in the layout k_update reads
// (permuted rows).  On the FP32 path the system is the FP64 one: every A
// entry summed as k_assemble sums it (linear part, FP64 device conductances
// in term order, the gmin diagonal) and the FP64 Newton residual, for the
// correction; rounding an ill-conditioned Jacobian to float alone moves its
// Newton step.  count_fb: add one to the instance's fallback count.
// Dynamic LDS: DENSE_LDS bytes (n <= 4096).
__device__ double slot_value64(const DevTopo& T, const DevBatch& P, int b, int q, double gmin) {
#pragma clang fp contract(off)
    const int n = T.n, Bp = P.Bp;
    auto V = [&](int node) -> double { return (node > 0 && node <= n) ? P.vI[IX(node - 1, b, Bp)] : 0.0; };
    double a = P.glin[IX(q, b, Bp)];
    for (int k = T.nl_ptr[q]; k < T.nl_ptr[q + 1]; ++k) {
        const int t = T.nl_term[k];
        const int kind = term_kind(t), i = term_idx(t);
        double g;
        if (kind == karam::SK_DIODE) {
            double id;
            diode_eval(V(T.d_a[i]) - V(T.d_c[i]), P.IsI[IX(i, b, Bp)], T.d_n[i], diode_vt(T, P, i, b), id, g);
        } else {
            double vgs, vds, ids, gm, gds;
            mos_eval(V(T.m_g[i]), V(T.m_d[i]), V(T.m_s[i]), T.m_w[i], T.m_l[i], P.kpI[IX(i, b, Bp)],
                     P.vthI[IX(i, b, Bp)], T.m_lam[i], T.m_pmos[i], vgs, vds, ids, gm, gds);
            g = kind == karam::SK_MOS_GDS ? gds : gm;
        }
        a = term_neg(t) ? a + (-g) : a + g;
    }
    if (T.slot_diag[q]) a = a + gmin;
    return a;
}

template <typename V>
__global__ __launch_bounds__(DENSE_THREADS) void k_dense(DevTopo T, DevBatch P, int count_fb,
                                                         const double* __restrict__ gmin_tab, int gmin_last) {
#pragma clang fp contract(off)
    extern __shared__ double dsm[];
    __shared__ int s_q, s_k;
    __shared__ double s_v;
    __shared__ double rv[DENSE_THREADS / 64];
    __shared__ int rk[DENSE_THREADS / 64];
    __shared__ int spiv[DENSE_NB_MAX];
    const int n = T.n, tid = threadIdx.x, nt = blockDim.x, Bp = P.Bp;
    const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
    double* sy = dsm;              // rhs -> y
    double* pan = dsm + n;         // panel (factor); products (backward)
    double* srow = pan;
    double* sl = pan + n;          // x (backward)
    const int cap = (int)(DENSE_LDS / sizeof(double)) - n;     // panel values
    double* M = P.dn_mat + (size_t)blockIdx.x * (size_t)n * (size_t)n;
    const V* pool = vp<V>(P.pool);
    const V* Al = vp<V>(P.Al);
    const V* rl = vp<V>(P.rhsl);
    for (;;) {
        if (tid == 0) s_q = atomicAdd(&P.dn_ctl[1], 1);
        __syncthreads();
        const int q = s_q;
        const int cnt = P.dn_ctl[0];
        __syncthreads();
        if (q >= cnt) return;
        const int b = P.dn_list[q];
        for (size_t e = tid; e < (size_t)n * (size_t)n; e += nt) M[e] = 0.0;
        __syncthreads();
        if (P.p32) {
            const double gmin = gmin_tab[min(P.st[b].step, gmin_last)];
            for (int i = tid; i < n; i += nt) {
                for (int k = T.row_ptr[i]; k < T.row_ptr[i + 1]; ++k)
                    M[(size_t)i * n + T.col_idx[k]] = 0.0 + slot_value64(T, P, b, k, gmin);
                sy[i] = P.rhslin[IX(i, b, Bp)] - gmin * P.vI[IX(i, b, Bp)];
            }
        } else {
            for (int i = tid; i < n; i += nt) {
                for (int k = T.row_ptr[i]; k < T.row_ptr[i + 1]; ++k) {
                    const int d = T.slot_dst[k];
                    const double a = (double)((d & 1) ? Al[IX(d >> 1, b, Bp)] : pool[IX(d >> 1, b, Bp)]);
                    M[(size_t)i * n + T.col_idx[k]] = 0.0 + a;
                }
                const int dn = T.node_dst[i];
                sy[i] = (double)((dn & 1) ? rl[IX(dn >> 1, b, Bp)] : pool[IX(dn >> 1, b, Bp)]);
            }
        }
        __syncthreads();
        int sing = 0;
        for (int i0 = 0, w = 0; i0 < n && !sing; i0 += w) {
            const int rows = n - i0;
            w = min(DENSE_NB_MAX, min(rows, cap / rows));
            // the panel: rows i0.., columns i0..i0+w-1, row-major (stride w)
            for (int e = tid; e < rows * w; e += nt) {
                const int r = e / w, k = e - r * w;
                pan[e] = M[(size_t)(i0 + r) * n + i0 + k];
            }
            __syncthreads();
            for (int k = 0; k < w; ++k) {
                double bv = 0.0;
                int bk = 0x7fffffff;
                for (int r = k + tid; r < rows; r += nt) {
                    const double v = fabs(pan[r * w + k]);
                    if (v > bv) { bv = v; bk = r; }
                }
                for (int o = 32; o > 0; o >>= 1) piv_better(bv, bk, __shfl_xor(bv, o), __shfl_xor(bk, o));
                if (lane == 0) { rv[wv] = bv; rk[wv] = bk; }
                __syncthreads();
                if (tid == 0) {
                    double v = rv[0];
                    int kk = rk[0];
                    for (int u = 1; u < nw; ++u) piv_better(v, kk, rv[u], rk[u]);
                    s_v = v;
                    s_k = v > 0.0 ? kk : k;
                    spiv[k] = i0 + s_k;
                }
                __syncthreads();
                if (s_v < 1e-25) { sing = 1; break; }
                const int p = s_k;
                if (p != k) {
                    for (int c = tid; c < w; c += nt) {
                        const double tmp = pan[k * w + c];
                        pan[k * w + c] = pan[p * w + c];
                        pan[p * w + c] = tmp;
                    }
                    if (tid == 0) { const double tb = sy[i0 + k]; sy[i0 + k] = sy[i0 + p]; sy[i0 + p] = tb; }
                    __syncthreads();
                }
                const double piv = pan[k * w + k];
                for (int r = k + 1 + tid; r < rows; r += nt) pan[r * w + k] = pan[r * w + k] / piv;
                __syncthreads();
                if (k + 1 < w) {
                    for (int r = k + 1 + tid; r < rows; r += nt) {
                        const double l = pan[r * w + k];
                        if (l != 0.0)
                            for (int c = k + 1; c < w; ++c) pan[r * w + c] = pan[r * w + c] - l * pan[k * w + c];
                    }
                    __syncthreads();
                }
            }
            if (sing) break;
            for (int e = tid; e < rows * w; e += nt) {
                const int r = e / w, k = e - r * w;
                M[(size_t)(i0 + r) * n + i0 + k] = pan[e];
            }
            // the panel's row swaps on every other column, in step order
            for (int cc = tid; cc < n - w; cc += nt) {
                const int c = cc < i0 ? cc : cc + w;
                for (int k = 0; k < w; ++k) {
                    const int q = spiv[k];
                    if (q != i0 + k) {
                        const double a = M[(size_t)(i0 + k) * n + c], bq = M[(size_t)q * n + c];
                        M[(size_t)(i0 + k) * n + c] = bq;
                        M[(size_t)q * n + c] = a;
                    }
                }
            }
            __syncthreads();
            // right of the panel: the pivot rows (U12), then the trailing
            // rows; a column's trailing rows may be split over S segments
            // (each recomputes the column's U12 in registers, segment 0
            // stores it)
            const int c0 = i0 + w, nh = n - c0, nr = rows - w;
            if (nh > 0) {
                const int S = max(1, nt / nh);
                const int len = (nr + S - 1) / S;
                for (int it = tid; it < nh * S; it += nt) {
                    const int c = c0 + it % nh, seg = it / nh;
                    double u[DENSE_NB_MAX];
#pragma unroll
                    for (int k = 0; k < DENSE_NB_MAX; ++k) u[k] = k < w ? M[(size_t)(i0 + k) * n + c] : 0.0;
#pragma unroll
                    for (int k = 1; k < DENSE_NB_MAX; ++k) {
                        if (k < w) {
                            double a = u[k];
#pragma unroll
                            for (int q = 0; q < k; ++q) {
                                const double l = pan[k * w + q];
                                if (l != 0.0) a = a - l * u[q];
                            }
                            u[k] = a;
                        }
                    }
                    if (seg == 0)
#pragma unroll
                        for (int k = 1; k < DENSE_NB_MAX; ++k)
                            if (k < w) M[(size_t)(i0 + k) * n + c] = u[k];
                    const int r0 = w + seg * len, r1 = min(rows, w + (seg + 1) * len);
                    int r = r0;
                    // with every u finite, a - l * u equals a when l is zero
                    // (the matrix holds no -0: entries start as 0.0 + a and a
                    // difference is -0 only from -0), so the updates need
                    // no test
                    bool fin = true;
#pragma unroll
                    for (int k = 0; k < DENSE_NB_MAX; ++k) fin = fin && __builtin_isfinite(u[k]);
                    if (fin) {
                        // DENSE_ROWS rows' loads in flight per lane
                        for (; r + DENSE_ROWS - 1 < r1; r += DENSE_ROWS) {
                            double a[DENSE_ROWS];
#pragma unroll
                            for (int t = 0; t < DENSE_ROWS; ++t) a[t] = M[(size_t)(i0 + r + t) * n + c];
#pragma unroll
                            for (int k = 0; k < DENSE_NB_MAX; ++k) {
                                if (k < w) {
#pragma unroll
                                    for (int t = 0; t < DENSE_ROWS; ++t) a[t] = a[t] - pan[(r + t) * w + k] * u[k];
                                }
                            }
#pragma unroll
                            for (int t = 0; t < DENSE_ROWS; ++t) M[(size_t)(i0 + r + t) * n + c] = a[t];
                        }
                    }
                    for (; r + 3 < r1; r += 4) {
                        double a[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t) a[t] = M[(size_t)(i0 + r + t) * n + c];
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const double* lr = pan + (r + t) * w;
#pragma unroll
                            for (int k = 0; k < DENSE_NB_MAX; ++k) {
                                if (k < w) {
                                    const double l = lr[k];
                                    if (l != 0.0) a[t] = a[t] - l * u[k];
                                }
                            }
                        }
#pragma unroll
                        for (int t = 0; t < 4; ++t) M[(size_t)(i0 + r + t) * n + c] = a[t];
                    }
                    for (; r < r1; ++r) {
                        const double* lr = pan + r * w;
                        double a = M[(size_t)(i0 + r) * n + c];
#pragma unroll
                        for (int k = 0; k < DENSE_NB_MAX; ++k) {
                            if (k < w) {
                                const double l = lr[k];
                                if (l != 0.0) a = a - l * u[k];
                            }
                        }
                        M[(size_t)(i0 + r) * n + c] = a;
                    }
                }
            }
            __syncthreads();
        }
        if (!sing) {
            for (int j = 0; j < n; ++j) {
                const double yj = sy[j];
                for (int i = j + 1 + tid; i < n; i += nt) sy[i] = sy[i] - M[(size_t)i * n + j] * yj;
                __syncthreads();
            }
            for (int i = n - 1; i >= 0; --i) {
                for (int j = i + 1 + tid; j < n; j += nt) srow[j] = M[(size_t)i * n + j] * sl[j];
                __syncthreads();
                if (tid == 0) {
                    double xi = sy[i];
                    for (int j = i + 1; j < n; ++j) xi = xi - srow[j];
                    sl[i] = xi / M[(size_t)i * n + i];
                }
                __syncthreads();
            }
        }
        double* xm = P.xm + (size_t)b * n;
        for (int i = tid; i < n; i += nt) xm[T.iperm[i]] = sing ? 0.0 : sl[i];
        if (tid == 0) {
            P.st[b].singular = sing;
            if (count_fb) P.st[b].lu_fb += 1;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// kernels: NR bookkeeping
// ---------------------------------------------------------------------------
// Per-instance NR control after a linear solve (one block per instance).
// tran = 1: a transient attempt (no limiting, no damping, KCL_TOL = 1e-12;
// solvers/circuitsim_solver.cpp:600-635 restated); a converged or failed
// attempt leaves the active list and k_tran_control decides.  On the FP32 path the solve returned the correction,
// so the new iterate is v + x (FP64), and every KCL audit also stores its
// per-node vector (the next Newton rhs).
// one ConvergenceStep of instance b (solvers/circuitsim_solver.cpp:407-441):
// the damping factor reads 0.5 when the (damped) residual is still not below
// the last one; the worst node as the reference fills it, from a vector that
// holds the residual at every node (so node 1); stagnation when the step's
// iteration is past the first and the residual moved by less than 1e-15
// above KCL_TOL (1e-12, physics/physics_constants.h:76); divergence when the
// last three residuals rise
__device__ void diag_record(const DevBatch& P, int b, const InstState& st, double kcl, double dv, double prev) {
    DiagState d = P.diag[b];
    const double damping = (st.it > 0 && kcl >= prev && st.polished == 0) ? 0.5 : 1.0;
    if (P.trace_cap > 0) {
        ConvRec r;
        r.residual = kcl; r.max_dv = dv; r.damping = damping;
        r.gmin_step = st.step; r.iteration = st.it; r.worst_node = 1; r.pad = 0;
        P.trace[(size_t)b * P.trace_cap + (size_t)(d.n % P.trace_cap)] = r;
    }
    if (st.it > 0 && fabs(prev - kcl) < 1e-15 && kcl > 1e-12) d.stagnation = 1;
    if (dv > 100.0) d.big_dv = 1;
    d.last[d.n % 10] = kcl;
    d.n++;
    if (d.n >= 3) {
        const double h1 = d.last[(d.n - 1) % 10], h2 = d.last[(d.n - 2) % 10], h3 = d.last[(d.n - 3) % 10];
        if (h1 > h2 && h2 > h3) d.divergence = 1;
        const double pd = h2 - h3, cd = h1 - h2;
        if ((pd > 0 && cd < 0) || (pd < 0 && cd > 0)) d.changes++;
    }
    P.diag[b] = d;
}

// NT threads per instance: 256, or 1024 when few instances are active (the
// audits' loops over devices and nodes stride by the block; every node's
// sum is one thread's, a hub's is thread 0's in list order, so the results
// do not depend on NT)
template <int NT>
__global__ __launch_bounds__(NT) void k_update(DevTopo T, DevBatch P, const int* __restrict__ act, int max_iter,
                                               int nsteps, int tran) {
#pragma clang fp contract(off)
    __shared__ double red[NT / 64];
    __shared__ int redi[NT / 64];
    if ((int)blockIdx.x >= *P.nact) return;
    const int b = act[blockIdx.x];
    const int n = T.n;
    InstState st = P.st[b];
    if (st.singular && P.promote) {
        // the float solve was rejected: the iterate stays, and the iteration
        // is taken again on the FP64 path
        if (threadIdx.x == 0) {
            st.singular = 0;
            st.lu_fb++;
            st.done = DONE_PROMOTE;
            P.st[b] = st;
        }
        return;
    }
    if (st.singular) {
        // singular Jacobian: abort the GMIN step and the solve (the reference's count)
        if (threadIdx.x == 0) {
            st.total += st.it + 1;
            st.ref += 2 * (st.it + 1);
            st.prev_kcl = 1e9;
            if (tran) st.kcl = 1e9;
            st.status = NR_STATUS_SINGULAR;
            st.done = tran ? DONE_FAILED : DONE_FINISHED;
            P.st[b] = st;
        }
        return;
    }
    double* v = P.v + (size_t)b * n;
    double* vn = P.vn + (size_t)b * n;
    const double* xm = P.xm + (size_t)b * n;
    double* res = P.p32 ? P.res + (size_t)b * n : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double val = P.p32 ? v[i] + xm[T.iperm[i]] : xm[T.iperm[i]];
        if (isnan(val) || isinf(val)) val = 0.0;   // sanitize_vector
        vn[i] = val;
    }
    __syncthreads();
    // pnjlim in SoA order; dependency levels preserve the sequential result
    const double* Isb = P.Is + (size_t)b * T.nD;
    for (int l = 0; !tran && l < T.n_pn_levels; ++l) {
        for (int k = T.pn_level_ptr[l] + threadIdx.x; k < T.pn_level_ptr[l + 1]; k += blockDim.x) {
            const int i = T.pn_diode[k];
            const double Vt = P.pnvt ? P.pnvt[(size_t)b * T.nD + i] : T.d_pn_vt[i];
            const double Is = Isb[i] * (P.pnisf ? P.pnisf[(size_t)b * T.nD + i] : T.d_pn_isf[i]);
            const int p = T.d_a[i] - 1, q = T.d_c[i] - 1;
            const double vcrit = Vt * log(Vt / (sqrt(2.0) * Is));
            const double vjo = (p >= 0 ? v[p] : 0.0) - (q >= 0 ? v[q] : 0.0);
            const double vjn = (p >= 0 ? vn[p] : 0.0) - (q >= 0 ? vn[q] : 0.0);
            const double vl = pnjlim(vjn, vjo, Vt, vcrit);
            const double corr = vjn - vl;
            if (fabs(corr) > 1e-12) {
                if (p >= 0 && q >= 0) { vn[p] -= corr * 0.5; vn[q] += corr * 0.5; }
                else if (p >= 0) vn[p] -= corr;
                else if (q >= 0) vn[q] += corr;
            }
        }
        __syncthreads();
    }
    double kcl = block_kcl(T, P, b, vn, red, res);
    const double prev_final = st.prev_kcl;
    // a polish iteration (after the last GMIN step converged) takes the full
    // Newton step: at the root the max-norm KCL has reached its gmin floor and
    // no longer falls, while the voltages still move
    if (!tran && st.it > 0 && kcl >= st.prev_kcl && st.polished == 0) {
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) vn[i] = v[i] + 0.5 * (vn[i] - v[i]);
            __syncthreads();
            kcl = block_kcl(T, P, b, vn, red, res);
            if (!(kcl >= st.prev_kcl)) break;
        }
    }
    st.prev_kcl = kcl;
    const double tol = (tran || st.step == nsteps - 1) ? 1e-12 : 1.0;
    int conv;
    if (kcl < 1e-9) {
        conv = 1;
    } else {
        int ok = 1;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const double lim = 1e-3 * ref_max(fabs(vn[i]), fabs(v[i])) + 1e-6;
            if (fabs(vn[i] - v[i]) > lim) ok = 0;
        }
        ok = block_and(ok, redi);
        conv = ok && !(kcl > tol);
    }
    double dv = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = fabs(vn[i] - v[i]);
        if (d > dv) dv = d;
        v[i] = vn[i];
    }
    if (!tran) {
        dv = block_max(dv, red);
        if (threadIdx.x == 0) diag_record(P, b, st, kcl, dv, prev_final);
    }
    if (threadIdx.x == 0 && tran) {
        st.kcl = kcl;                         // the last iterate's audit
        if (conv) {
            st.total += st.it + 1;
            st.ref += st.it + 1;
            st.done = DONE_CONVERGED;         // the step control decides
        } else if (++st.it >= max_iter) {
            st.total += max_iter;
            st.ref += max_iter;
            st.status = NR_STATUS_NOT_CONVERGED;
            st.done = DONE_FAILED;
        }
        P.st[b] = st;
    } else if (threadIdx.x == 0) {
        if (conv && st.step == nsteps - 1 && st.polished < P.polish && st.it + 1 < max_iter) {
            // one more Newton iteration from the converged iterate
            st.polished++;
            st.it++;
        } else if (conv) {
            st.total += st.it + 1;
            st.ref += st.it + 1;
            if (st.step == nsteps - 1) { st.status = NR_STATUS_CONVERGED; st.done = 1; }
            else { st.step++; st.it = 0; }
        } else {
            st.it++;
            if (st.it >= max_iter) {
                st.total += max_iter;
                st.ref += max_iter + 1;
                if (st.step == nsteps - 1 || !(st.prev_kcl < 1.0)) {
                    st.status = NR_STATUS_NOT_CONVERGED;
                    st.done = 1;
                } else {
                    st.step++;
                    st.it = 0;
                }
            }
        }
        P.st[b] = st;
    }
}

// The instances that left the FP32 sweeps resume on the FP64 path (one
// block per instance; P is the FP64 view): the interleaved rhs goes back to
// the linear rhs (the FP32 path keeps the KCL vector there)
__global__ __launch_bounds__(256) void k_promote(DevTopo T, DevBatch P) {
    const int b = blockIdx.x;
    if (P.st[b].done != DONE_PROMOTE) return;
    for (int i = threadIdx.x; i < T.n; i += blockDim.x) P.rhslin[IX(i, b, P.Bp)] = rhs_lin_node(T, P, b, i);
    __syncthreads();
    if (threadIdx.x == 0) P.st[b].done = DONE_RUNNING;
}

// rebuild the active list in increasing instance order (single block)
__global__ __launch_bounds__(1024) void k_compact(DevBatch P) {
    __shared__ int cnt[1024];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int per = (P.B + nt - 1) / nt;
    const int lo = tid * per, hi = min(P.B, lo + per);
    int c = 0;
    for (int i = lo; i < hi; ++i) c += P.st[i].done ? 0 : 1;
    cnt[tid] = c;
    __syncthreads();
    for (int o = 1; o < nt; o <<= 1) {
        int v = (tid >= o) ? cnt[tid - o] : 0;
        __syncthreads();
        cnt[tid] += v;
        __syncthreads();
    }
    int pos = cnt[tid] - c;
    for (int i = lo; i < hi; ++i)
        if (!P.st[i].done) P.act[pos++] = i;
    if (tid == nt - 1) *P.nact = cnt[tid];
}

// ---------------------------------------------------------------------------
// AC small-signal analysis (ACSolver::solveAC, solvers/ac_solver.h:339-383):
// Y(jw) = G + jwC at the operating point, solved per frequency by the same
// multifrontal kernels on complex values
// ---------------------------------------------------------------------------
// Device conductances at the operating point in vI (lane per (device,
// instance)).  Diodes as buildAdmittanceMatrix linearises them (:186-196):
// the derivative of diode_current_clamped as its dual numbers compute it
// (Is exp(arg) n_vt / n_vt^2, or (Is / n_vt) e^30 past arg = 30) with the
// netlist's unscaled thermal voltage, at least 1e-15.  MOSFETs: gm and gds of
// the generic kernel with the instance's parameters (the DC Jacobian's).
__global__ __launch_bounds__(256) void k_ac_eval(DevTopo T, DevBatch P, const int* __restrict__ act, int nact) {
#pragma clang fp contract(off)
    LANE_SETUP(4)
    const int n = T.n;
    auto V = [&](int node) -> double { return (node > 0 && node <= n) ? P.vI[IX(node - 1, b, Bp)] : 0.0; };
    if (item < T.nD) {
        const int i = item;
        const double vd = V(T.d_a[i]) - V(T.d_c[i]);
        const double Is = P.IsI[IX(i, b, Bp)], n_vt = T.d_n[i] * T.d_vtraw[i];
        const double arg = vd / n_vt;
        const double g = arg > 30.0 ? (Is / n_vt) * exp(30.0) : Is * (exp(arg) * (n_vt / (n_vt * n_vt)));
        P.dg[IX(i, b, Bp)] = g < 1e-15 ? 1e-15 : g;
    } else if (item < T.nD + T.nM) {
        const int j = item - T.nD;
        const int d = T.m_d[j], g = T.m_g[j], s = T.m_s[j];
        double vgs, vds, ids, gm, gds;
        mos_eval(V(g), V(d), V(s), T.m_w[j], T.m_l[j], P.kpI[IX(j, b, Bp)], P.vthI[IX(j, b, Bp)], T.m_lam[j],
                 T.m_pmos[j], vgs, vds, ids, gm, gds);
        P.mgds[IX(j, b, Bp)] = gds;
        P.mgm[IX(j, b, Bp)] = gm;
    }
}

// Complex A and rhs of instance b at angular frequency omega into Q's
// factor inputs (lane per (slot, instance)).  Each slot sums its stamps in
// emission order, real and imaginary parts alike: resistors 1/R, Norton
// V-source conductances 1e3 and capacitors j omega C (the linear stamps),
// then the device conductances, the 1e-12 diagonal and, at the excited
// node, 1e6 (:364-375); the excited node's rhs is 1e6 x amplitude.
__global__ __launch_bounds__(256) void k_ac_assemble(DevTopo T, DevBatch P, DevBatch Q, const int* __restrict__ act,
                                                     int nact, double omega, int in_slot, int in_node,
                                                     double g_large, double amplitude) {
#pragma clang fp contract(off)
    LANE_SETUP(4)
    if (item < T.nnz_a) {
        double re = 0.0, im = 0.0;
        for (int k = T.lin_ptr[item]; k < T.lin_ptr[item + 1]; ++k) {
            const int t = T.lin_term[k];
            const int idx = term_idx(t);
            double gr = 0.0, gi = 0.0;
            switch (term_kind(t)) {
            case karam::SK_RES: gr = 1.0 / P.R[(size_t)b * T.nR + idx]; break;
            case karam::SK_CAP: gi = omega * P.cap[(size_t)b * T.nC + idx]; break;
            default: gr = 1.0e3; break;
            }
            if (term_neg(t)) { re = re + (-gr); im = im + (-gi); }
            else { re = re + gr; im = im + gi; }
        }
        for (int k = T.nl_ptr[item]; k < T.nl_ptr[item + 1]; ++k) {
            const int t = T.nl_term[k];
            const int kind = term_kind(t);
            const double* src = kind == karam::SK_DIODE ? P.dg : kind == karam::SK_MOS_GDS ? P.mgds : P.mgm;
            const double g = src[IX(term_idx(t), b, Bp)];
            if (term_neg(t)) { re = re + (-g); im = im + (-0.0); }
            else { re = re + g; im = im + 0.0; }
        }
        if (T.slot_diag[item]) {
            re = re + 1e-12;
            im = im + 0.0;
            if (item == in_slot) {
                re = re + g_large;
                im = im + 0.0;
            }
        }
        const int d = T.slot_dst[item];
        const cdouble a(re, im);
        if (d & 1) vp<cdouble>(Q.Al)[IX(d >> 1, b, Bp)] = a;
        else vp<cdouble>(Q.pool)[IX(d >> 1, b, Bp)] = a;
    }
    if (item < T.n) {
        const cdouble r = item == in_node - 1 ? cdouble(0.0 + g_large * amplitude, 0.0 + 0.0) : cdouble(0.0, 0.0);
        const int d = T.node_dst[item];
        if (d & 1) vp<cdouble>(Q.rhsl)[IX(d >> 1, b, Bp)] = r;
        else vp<cdouble>(Q.pool)[IX(d >> 1, b, Bp)] = r;
    }
}

// clear the complex pass's singular flags (thread per instance)
__global__ void k_ac_reset(DevBatch Q, int nact) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nact) Q.st[b].singular = 0;
}

// solution at the probed nodes of frequency point f -> out[b][f][p] (re, im);
// singular instances are flagged in status[b]
__global__ void k_ac_out(DevTopo T, DevBatch Q, int nact, int f, int n_freq, int n_probe,
                         const int* __restrict__ probe, double* __restrict__ out, int* __restrict__ status) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long long)nact * n_probe) return;
    const int b = (int)(k / n_probe), p = (int)(k % n_probe);
    const cdouble x = vp<cdouble>(Q.xs)[IX(T.iperm[probe[p] - 1], b, Q.Bp)];
    double* o = out + (((size_t)b * n_freq + f) * n_probe + p) * 2;
    o[0] = x.re;
    o[1] = x.im;
    if (p == 0 && Q.st[b].singular) status[b] = NR_STATUS_SINGULAR;
}

// Noise sources (solvers/ac_solver.h:400-504) in computeNoiseSources order:
// src[k] = (kind, device, pos, neg) with kind 0 a resistor (netlist index;
// 4kT/R), 1 a diode (SoA index; 2q|I| of the clamped diode current at the
// operating point), 2 a MOSFET (SoA index; 4kT (2/3) gm when gm > 0).  Their
// PSDs do not depend on frequency: k_noise_psd writes them once per run,
// interleaved (psd[IX(k, b)]), with +0 for a source whose PSD is not positive
// (the reference leaves it out of its list, physics/noise_models.h; +0 leaves
// the sum of non-negative terms unchanged).
constexpr double NOISE_KB = 1.380649e-23, NOISE_QE = 1.602176634e-19;
__global__ __launch_bounds__(256) void k_noise_psd(DevTopo T, DevBatch P, int nact, const int4* __restrict__ src,
                                                   int n_src, double temp_K, double* __restrict__ psd) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int Bp = P.Bp;
    if (t >= (long long)n_src * Bp) return;
    const int k = (int)(t / Bp), b = (int)(t % Bp);
    if (b >= nact) return;
    const int n = T.n;
    auto V = [&](int node) -> double { return (node > 0 && node <= n) ? P.vI[IX(node - 1, b, Bp)] : 0.0; };
    const int4 e = src[k];
    double s;
    if (e.x == 0) {
        const double R = P.R[(size_t)b * T.nR + e.y];
        s = R <= 0.0 ? 0.0 : 4.0 * NOISE_KB * temp_K / R;
    } else if (e.x == 1) {
        const double I = diode_i_clamped(V(e.z) - V(e.w), P.Is[(size_t)b * T.nD + e.y], T.d_n[e.y], T.d_vtraw[e.y]);
        s = 2.0 * NOISE_QE * fabs(I);
    } else {
        const double gm = P.mgm[IX(e.y, b, Bp)];
        s = gm <= 0.0 ? 0.0 : 4.0 * NOISE_KB * temp_K * (2.0 / 3.0) * gm;
    }
    psd[IX(k, b, Bp)] = s > 0.0 ? s : 0.0;
}

// Noise of frequency point f: after the solve of Y^T z = e_out (Q.xs), the
// output noise PSD sum_k |z[pos_k] - z[neg_k]|^2 S_k in source order.
// pn[k] = (pos, neg) as rows of the permuted solution (-1 for ground).  One
// block per 64 instances: its sixteen wavefronts evaluate 256 terms at a time
// into LDS, and wavefront 0 adds them in source order.
__device__ __forceinline__ double noise_term(const DevBatch& Q, int b, int2 pn, double s) {
#pragma clang fp contract(off)
    const int Bp = Q.Bp;
    const cdouble zp = pn.x >= 0 ? vp<cdouble>(Q.xs)[IX(pn.x, b, Bp)] : cdouble(0.0, 0.0);
    const cdouble zn = pn.y >= 0 ? vp<cdouble>(Q.xs)[IX(pn.y, b, Bp)] : cdouble(0.0, 0.0);
    const double hr = zp.re - zn.re, hi = zp.im - zn.im;
    return s > 0.0 ? (hr * hr + hi * hi) * s : 0.0;
}

__global__ __launch_bounds__(HUB_SUM_THREADS) void k_noise_sum(DevBatch Q, int nact, const int2* __restrict__ pn,
                                                               const double* __restrict__ psd, int n_src, int f,
                                                               int n_freq, double* __restrict__ out,
                                                               int* __restrict__ status) {
#pragma clang fp contract(off)
    __shared__ double buf[HUB_SUM_CHUNK][64];
    constexpr int PER = HUB_SUM_CHUNK / (HUB_SUM_THREADS / 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b0 = blockIdx.x * 64, b = b0 + lane;
    const bool on = b < nact;
    const int bb = on ? b : b0;
    double total = 0.0;
    for (int base = 0; base < n_src; base += HUB_SUM_CHUNK) {
        const int cnt = min(HUB_SUM_CHUNK, n_src - base);
        double x[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = w * PER + u;
            x[u] = q < cnt ? noise_term(Q, bb, pn[base + q], psd[IX(base + q, bb, Q.Bp)]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) buf[w * PER + u][lane] = x[u];
        __syncthreads();
        if (w == 0)
            for (int q = 0; q < cnt; ++q) total = total + buf[q][lane];
        __syncthreads();
    }
    if (w == 0 && on) {
        out[(size_t)b * n_freq + f] = total;
        if (Q.st[b].singular) status[b] = NR_STATUS_SINGULAR;
    }
}

// sqrt of the trapezoid integral of each instance's PSD over the frequency
// points (solvers/ac_solver.h:585-592)
__global__ void k_noise_integrate(const double* __restrict__ freqs, const double* __restrict__ psd, int n_freq,
                                  int nact, double* __restrict__ vrms) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nact) return;
    const double* p = psd + (size_t)b * n_freq;
    double integrated = 0.0;
    for (int i = 1; i < n_freq; ++i) {
        const double df = freqs[i] - freqs[i - 1];
        const double avg = 0.5 * (p[i] + p[i - 1]);
        integrated += avg * df;
    }
    vrms[b] = sqrt(integrated > 0.0 ? integrated : 0.0);
}

// The final audit of each instance: the max |KCL| and, from the per-node
// residuals (staged in vn, free after the solve), their RMS (squares added
// in node order), the first node with the largest |residual| and its voltage
// (solvers/circuitsim_solver.cpp:489-508).
__global__ __launch_bounds__(256) void k_final_kcl(DevTopo T, DevBatch P) {
#pragma clang fp contract(off)
    __shared__ double red[8];
    __shared__ double wbest[8];
    __shared__ int widx[8];
    const int b = blockIdx.x, n = T.n;
    const double* v = P.v + (size_t)b * n;
    double* rr = P.vn + (size_t)b * n;
    double k = block_kcl(T, P, b, v, red, rr);
    __syncthreads();
    const double ss = block_ordered_sum(n, [&](int i) {
        const double a = fabs(rr[i]);
        return a * a;
    });
    double best = 0.0;
    int bi = -1;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double a = fabs(rr[i]);
        if (a > best) { best = a; bi = i; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_down(best, o);
        const int oi = __shfl_down(bi, o);
        if (oi >= 0 && (ob > best || (ob == best && (bi < 0 || oi < bi)))) { best = ob; bi = oi; }
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { wbest[w] = best; widx[w] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            const double ob = wbest[q];
            const int oi = widx[q];
            if (oi >= 0 && (ob > best || (ob == best && (bi < 0 || oi < bi)))) { best = ob; bi = oi; }
        }
        P.st[b].kcl = k;
        FinalDiag f;
        f.rms = sqrt(ss / (double)n);
        f.worst_node = bi + 1;
        f.worst_v = bi >= 0 ? v[bi] : 0.0;
        f.pad = 0;
        P.fin[b] = f;
    }
}

// standalone evaluation slots
__global__ void k_eval_diodes_slot(int count, const int* a, const int* c, const double* Is, const double* N,
                                   const double* Vt, const double* v, int nv, double* vd, double* id, double* gd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double x = volt(v, nv, a[i]) - volt(v, nv, c[i]);
    double ii, g;
    diode_eval(x, Is[i], N[i], Vt[i], ii, g);
    vd[i] = x; id[i] = ii; gd[i] = g;
}

__global__ void k_eval_mos_slot(int count, const int* d, const int* g, const int* s, const double* W,
                                const double* L, const double* Kp, const double* Vth, const double* lam,
                                const int* pm, const double* v, int nv, double* vgs, double* vds, double* ids,
                                double* gm, double* gds) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    mos_eval(volt(v, nv, g[i]), volt(v, nv, d[i]), volt(v, nv, s[i]), W[i], L[i], Kp[i], Vth[i], lam[i], pm[i],
             vgs[i], vds[i], ids[i], gm[i], gds[i]);
}

// Staging of the no-allocation MOSFET evaluation slot (nr_eval_mosfets_noalloc):
// one call's inputs and outputs.  Every GPU the library runs on holds its own
// copy until the process ends.  Host code reaches it only through the address
// hipGetSymbolAddress returns on each device (kept per device by
// nr_ctx_create), never through this name, which on the host is not a device
// address.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

#include "nr_host.inc"
#include "seeder.inc"
