// nr_device.hip — batched Newton-Raphson DC solve on MI355X (gfx950).
//
// One batch = B Monte-Carlo / corner instances of one topology.  Per batched
// NR sweep over the still-active instances:
//   k_eval        device evaluation, one lane per (instance, device)
//                 (generic kernels, tensors/physics_tensors.cpp:19-104)
//   k_assemble    A and rhs through the precomputed stamp->slot lists, each
//                 slot summed in the reference's stamp order (assembleFast,
//                 math/linalg.h:441-465) — no atomics
//   k_factor      multifrontal LU of the permuted system on the reference's
//                 Min-Fill fill pattern, one team per (supernode, instance),
//                 fronts staged in LDS, forward elimination fused; level by
//                 level of the supernodal elimination tree
//   k_backsolve   back substitution, levels top-down
//   k_update      per instance: unpermute, NaN guard, pnjlim, FP64 KCL
//                 residual, damping, convergence test, GMIN-step control
//                 (solvers/circuitsim_solver.cpp:176-555)
//   k_compact     retire finished instances from the active list
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "karam_nr.h"
#include "topology.h"

using karam::Topology;

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static thread_local std::string g_lib_error;

struct nr_ctx {
    int device = 0;
    int verbose = 0;
    hipStream_t stream = nullptr;
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
struct DevTopo {
    int n, nnz_a, nD, nM, nR, nV, nI;
    const int *d_a, *d_c;
    const double *d_n, *d_vt, *d_vtraw, *d_pn_vt, *d_pn_isf;
    const int *m_d, *m_g, *m_s, *m_b, *m_pmos;
    const double *m_w, *m_l, *m_lam;
    const int *res_n1, *res_n2, *vsrc_pos, *vsrc_neg, *isrc_pos, *isrc_neg;
    const int *dio_anode_nl;            // netlist-order anodes (seed)
    const int *lin_ptr, *lin_term, *nl_ptr, *nl_term, *slot_diag;
    const int *rhs_lin_ptr, *rhs_lin_term, *rhs_nl_ptr, *rhs_nl_term;
    const int *kcl_ptr, *kcl_term;
    const int *pn_level_ptr, *pn_diode;
    int n_pn_levels;
    const int *perm, *iperm;
    const int *sn_first, *sn_ncols, *sn_nrows, *sn_row_off, *sn_a_begin, *sn_a_end;
    const int *sn_child_begin, *sn_child_end;
    const long long *sn_u_off, *sn_cb_off, *sn_rcb_off;
    const int *children, *row_idx, *relmap, *fa_slot, *fa_li, *fa_lj;
    long long u_size, cb_size, rcb_size;
};

struct InstState {
    int step, it, total, ref;
    int status, done, singular, seeded;
    double prev_kcl, kcl;
};

struct DevBatch {
    int B;
    // per-instance parameters [B][count], netlist order (Is in SoA order)
    double *R, *Is, *kp, *vth, *isrc, *vsrc;
    // state
    double *v, *vn, *glin, *rhslin, *A, *rhs;
    double *dg, *dieq, *mgds, *mgm, *mieq;
    double *U, *y, *x, *CB, *RCB;
    double *seed;                     // optional [B][n]
    InstState* st;
    int* act;
    int* nact;
};

// ---------------------------------------------------------------------------
// device helpers (reference arithmetic; no FP contraction so the evaluation
// matches the CPU restatement operation for operation)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double volt(const double* v, int n, int node) {
    return (node > 0 && node <= n) ? v[node - 1] : 0.0;
}

// batchDiodePhysics_generic (tensors/physics_tensors.cpp:19-51)
__device__ __forceinline__ void diode_eval(double vd, double Is, double N, double Vt,
                                           double& i, double& g) {
#pragma clang fp contract(off)
    double n_vt = N * Vt;
    double vcrit = 30.0 * n_vt;
    double arg = vd / n_vt;
    double argc = (30.0 < arg) ? 30.0 : arg;          // std::min(arg, 30.0)
    double e = exp(argc);
    double ib = Is * (e - 1.0);
    double gb = (Is / n_vt) * e;
    double d = vd - vcrit;
    double delta = (0.0 < d) ? d : 0.0;              // std::max(0.0, d)
    i = ib + delta * gb;
    g = (1e-12 < gb) ? gb : 1e-12;                   // std::max(GMIN, gb)
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
__device__ __forceinline__ void mos_eval(double vg, double vd_, double vs, double W, double L,
                                         double Kp, double Vth, double lam, int pmos,
                                         double& vgs, double& vds, double& ids, double& gm,
                                         double& gds) {
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
        vlim = (vnew < lo) ? lo : vnew;              // std::max(vnew, lo)
    }
    return vlim;
}

__device__ __forceinline__ int term_kind(int t) { return (t >> 1) & 7; }
__device__ __forceinline__ bool term_neg(int t) { return (t & 1) != 0; }
__device__ __forceinline__ int term_idx(int t) { return t >> 4; }

// block-wide reductions (256 threads)
__device__ __forceinline__ double block_max(double x, double* red) {
    for (int o = 32; o > 0; o >>= 1) {
        double y = __shfl_xor(x, o);
        x = (x < y || (y != y)) ? y : x;           // NaN propagates
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = x;
    __syncthreads();
    double r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
        double y = red[i];
        r = (r < y || (y != y)) ? y : r;
    }
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

// ---------------------------------------------------------------------------
// KCL residual of one node (calculatePhysicalResiduals terms,
// physics/circuitsim.h:1250-1671, D7-restated MOSFET term)
// ---------------------------------------------------------------------------
__device__ double kcl_node(const DevTopo& T, const DevBatch& P, int b, const double* v, int i) {
#pragma clang fp contract(off)
    const int n = T.n;
    double s = 0.0;
    for (int k = T.kcl_ptr[i]; k < T.kcl_ptr[i + 1]; ++k) {
        const int t = T.kcl_term[k];
        const int idx = term_idx(t);
        double val;
        switch (term_kind(t)) {
        case karam::TK_RES:
            val = (volt(v, n, T.res_n1[idx]) - volt(v, n, T.res_n2[idx])) / P.R[(size_t)b * T.nR + idx];
            break;
        case karam::TK_VSRC:
            val = (P.vsrc[(size_t)b * T.nV + idx] - (volt(v, n, T.vsrc_pos[idx]) - volt(v, n, T.vsrc_neg[idx]))) / 1e-3;
            break;
        case karam::TK_ISRC:
            val = P.isrc[(size_t)b * T.nI + idx];
            break;
        case karam::TK_DIODE: {
            double vd = volt(v, n, T.d_a[idx]) - volt(v, n, T.d_c[idx]);
            val = diode_i_clamped(vd, P.Is[(size_t)b * T.nD + idx], T.d_n[idx], T.d_vtraw[idx]);
            break;
        }
        default: {   // TK_MOS
            double vgs, vds, ids, gm, gds;
            mos_eval(volt(v, n, T.m_g[idx]), volt(v, n, T.m_d[idx]), volt(v, n, T.m_s[idx]),
                     T.m_w[idx], T.m_l[idx], P.kp[(size_t)b * T.nM + idx], P.vth[(size_t)b * T.nM + idx],
                     T.m_lam[idx], T.m_pmos[idx], vgs, vds, ids, gm, gds);
            val = ids;
            break;
        }
        }
        s = term_neg(t) ? s + (-val) : s + val;
    }
    return s;
}

__device__ double block_kcl(const DevTopo& T, const DevBatch& P, int b, const double* v, double* red) {
    double mx = 0.0;
    for (int i = threadIdx.x; i < T.n; i += blockDim.x) {
        double r = fabs(kcl_node(T, P, b, v, i));
        mx = (mx < r || (r != r)) ? r : mx;
    }
    return block_max(mx, red);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// seedInitialGuess (solvers/circuitsim_solver.cpp:32-116) or the caller's
// seed; then the per-instance linear part of A and rhs.
__global__ void k_init(DevTopo T, DevBatch P, int seeded_step) {
#pragma clang fp contract(off)
    const int b = blockIdx.x;
    const int n = T.n;
    double* v = P.v + (size_t)b * n;
    __shared__ int use_seed;
    if (threadIdx.x == 0) {
        int ok = 0;
        if (P.seed) {
            ok = 1;
            const double* sv = P.seed + (size_t)b * n;
            bool all_nan = true;
            for (int i = 0; i < n; ++i) if (sv[i] == sv[i]) { all_nan = false; break; }
            ok = all_nan ? 0 : 1;
        }
        use_seed = ok;
    }
    __syncthreads();
    if (use_seed) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = P.seed[(size_t)b * n + i];
    } else {
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
                if (a > 0 && a <= n) { double x = v[a - 1]; v[a - 1] = (x < 0.6) ? 0.6 : x; }
            }
            double vs = 5.0;
            for (int i = 0; i < T.nV; ++i) if (V[i] > vs) vs = V[i];
            for (int i = 0; i < T.nI; ++i) {
                int p = T.isrc_pos[i];
                if (p <= 0 || p > n) continue;
                if (v[p - 1] == 0.0) v[p - 1] = vs * 0.5;
            }
        }
    }
    // linear stamps
    for (int s = threadIdx.x; s < T.nnz_a; s += blockDim.x) {
        double a = 0.0;
        for (int k = T.lin_ptr[s]; k < T.lin_ptr[s + 1]; ++k) {
            const int t = T.lin_term[k];
            const int idx = term_idx(t);
            double g = (term_kind(t) == karam::SK_RES) ? 1.0 / P.R[(size_t)b * T.nR + idx] : 1.0e3;
            a = term_neg(t) ? a + (-g) : a + g;
        }
        P.glin[(size_t)b * T.nnz_a + s] = a;
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double r = 0.0;
        for (int k = T.rhs_lin_ptr[i]; k < T.rhs_lin_ptr[i + 1]; ++k) {
            const int t = T.rhs_lin_term[k];
            const int idx = term_idx(t);
            double val = (term_kind(t) == karam::TK_VSRC) ? P.vsrc[(size_t)b * T.nV + idx] * 1.0e3
                                                         : P.isrc[(size_t)b * T.nI + idx];
            r = term_neg(t) ? r + (-val) : r + val;
        }
        P.rhslin[(size_t)b * n + i] = r;
    }
    if (threadIdx.x == 0) {
        InstState s;
        s.step = use_seed ? seeded_step : 0;
        s.it = 0; s.total = 0; s.ref = 0;
        s.status = NR_STATUS_NOT_CONVERGED; s.done = 0; s.singular = 0; s.seeded = use_seed;
        s.prev_kcl = 0.0; s.kcl = 0.0;
        P.st[b] = s;
        P.act[b] = b;
    }
}

// device evaluation at the current estimate v
__global__ void k_eval(DevTopo T, DevBatch P, const int* __restrict__ act) {
#pragma clang fp contract(off)
    const int b = act[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = T.n;
    const double* v = P.v + (size_t)b * n;
    if (i < T.nD) {
        double vd = volt(v, n, T.d_a[i]) - volt(v, n, T.d_c[i]);
        double id, g;
        diode_eval(vd, P.Is[(size_t)b * T.nD + i], T.d_n[i], T.d_vt[i], id, g);
        P.dg[(size_t)b * T.nD + i] = g;
        P.dieq[(size_t)b * T.nD + i] = id - g * vd;
    } else if (i < T.nD + T.nM) {
        const int j = i - T.nD;
        const int d = T.m_d[j], g = T.m_g[j], s = T.m_s[j], bb = T.m_b[j];
        double vgs, vds, ids, gm, gds;
        mos_eval(volt(v, n, g), volt(v, n, d), volt(v, n, s), T.m_w[j], T.m_l[j],
                 P.kp[(size_t)b * T.nM + j], P.vth[(size_t)b * T.nM + j], T.m_lam[j], T.m_pmos[j],
                 vgs, vds, ids, gm, gds);
        const double gmb = 0.0;
        double vbs = (bb != 0) ? (v[bb - 1] - (s != 0 ? v[s - 1] : 0.0)) : 0.0;
        P.mgds[(size_t)b * T.nM + j] = gds;
        P.mgm[(size_t)b * T.nM + j] = gm;
        P.mieq[(size_t)b * T.nM + j] = ids - (gm * vgs + gmb * vbs + gds * vds);
    }
}

// A = linear part + device stamps (emission order) + gmin; rhs likewise
__global__ void k_assemble(DevTopo T, DevBatch P, const int* __restrict__ act, const double* __restrict__ gmin_tab) {
#pragma clang fp contract(off)
    const int b = act[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double gmin = gmin_tab[P.st[b].step];
    if (i < T.nnz_a) {
        double a = P.glin[(size_t)b * T.nnz_a + i];
        for (int k = T.nl_ptr[i]; k < T.nl_ptr[i + 1]; ++k) {
            const int t = T.nl_term[k];
            const int idx = term_idx(t);
            double g;
            switch (term_kind(t)) {
            case karam::SK_DIODE: g = P.dg[(size_t)b * T.nD + idx]; break;
            case karam::SK_MOS_GDS: g = P.mgds[(size_t)b * T.nM + idx]; break;
            default: g = P.mgm[(size_t)b * T.nM + idx]; break;
            }
            a = term_neg(t) ? a + (-g) : a + g;
        }
        if (T.slot_diag[i]) a = a + gmin;
        P.A[(size_t)b * T.nnz_a + i] = a;
    }
    if (i < T.n) {
        double r = P.rhslin[(size_t)b * T.n + i];
        for (int k = T.rhs_nl_ptr[i]; k < T.rhs_nl_ptr[i + 1]; ++k) {
            const int t = T.rhs_nl_term[k];
            const int idx = term_idx(t);
            double val = (term_kind(t) == karam::TK_DIODE) ? P.dieq[(size_t)b * T.nD + idx]
                                                          : P.mieq[(size_t)b * T.nM + idx];
            r = term_neg(t) ? r + (-val) : r + val;
        }
        P.rhs[(size_t)b * T.n + i] = r;
    }
}

template <int TEAM>
__device__ __forceinline__ void team_sync() {
    if constexpr (TEAM == 64) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    } else {
        __syncthreads();
    }
}

// Multifrontal partial factorisation of one supernode front per team.
template <int TEAM>
__global__ __launch_bounds__(256) void k_factor(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                int nact, const int* __restrict__ sn_list, int n_sn,
                                                int rmax) {
    extern __shared__ double smem[];
    constexpr int TPB = 256 / TEAM;
    const int tib = threadIdx.x / TEAM;
    const int lane = threadIdx.x % TEAM;
    const long long task = (long long)blockIdx.x * TPB + tib;
    if (task >= (long long)n_sn * nact) return;
    const int J = sn_list[task % n_sn];
    const int b = act[task / n_sn];
    const int c = T.sn_ncols[J], m = T.sn_nrows[J], r = c + m, first = T.sn_first[J];
    double* F = smem + (size_t)tib * (size_t)(rmax * rmax + rmax);
    double* fr = F + r * r;
    const double* A = P.A + (size_t)b * T.nnz_a;
    const double* rhs = P.rhs + (size_t)b * T.n;
    for (int e = lane; e < r * r + r; e += TEAM) F[e] = 0.0;
    team_sync<TEAM>();
    for (int e = T.sn_a_begin[J] + lane; e < T.sn_a_end[J]; e += TEAM)
        F[T.fa_li[e] * r + T.fa_lj[e]] = A[T.fa_slot[e]];
    for (int k = lane; k < c; k += TEAM) fr[k] = rhs[T.perm[first + k]];
    team_sync<TEAM>();
    for (int ch = T.sn_child_begin[J]; ch < T.sn_child_end[J]; ++ch) {
        const int C = T.children[ch];
        const int mc = T.sn_nrows[C];
        const int* rel = T.relmap + T.sn_row_off[C];
        const double* cb = P.CB + (size_t)b * T.cb_size + T.sn_cb_off[C];
        const double* rcb = P.RCB + (size_t)b * T.rcb_size + T.sn_rcb_off[C];
        for (int e = lane; e < mc * mc; e += TEAM) {
            const int a = e / mc, q = e - a * mc;
            F[rel[a] * r + rel[q]] += cb[e];
        }
        for (int a = lane; a < mc; a += TEAM) fr[rel[a]] += rcb[a];
        team_sync<TEAM>();
    }
    int sing = 0;
    for (int k = 0; k < c; ++k) {
        const double piv = F[k * r + k];
        if (!(fabs(piv) >= 1e-25)) sing = 1;
        for (int i = k + 1 + lane; i < r; i += TEAM) F[i * r + k] /= piv;
        team_sync<TEAM>();
        const int w = r - k - 1;
        const double fk = fr[k];
        for (int e = lane; e < w * w; e += TEAM) {
            const int i = k + 1 + e / w, j = k + 1 + e % w;
            F[i * r + j] -= F[i * r + k] * F[k * r + j];
        }
        for (int i = k + 1 + lane; i < r; i += TEAM) fr[i] -= F[i * r + k] * fk;
        team_sync<TEAM>();
    }
    if (sing && lane == 0) P.st[b].singular = 1;
    double* U = P.U + (size_t)b * T.u_size + T.sn_u_off[J];
    for (int e = lane; e < c * r; e += TEAM) U[e] = F[e];
    double* y = P.y + (size_t)b * T.n;
    for (int k = lane; k < c; k += TEAM) y[first + k] = fr[k];
    if (m > 0) {
        double* cb = P.CB + (size_t)b * T.cb_size + T.sn_cb_off[J];
        double* rcb = P.RCB + (size_t)b * T.rcb_size + T.sn_rcb_off[J];
        for (int e = lane; e < m * m; e += TEAM) {
            const int a = e / m, q = e - a * m;
            cb[e] = F[(c + a) * r + c + q];
        }
        for (int a = lane; a < m; a += TEAM) rcb[a] = fr[c + a];
    }
}

// Back substitution of one supernode per team: x_J = U11^{-1}(y_J - U12 x_R)
template <int TEAM>
__global__ __launch_bounds__(256) void k_backsolve(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                   int nact, const int* __restrict__ sn_list, int n_sn,
                                                   int rmax) {
    extern __shared__ double smem[];
    constexpr int TPB = 256 / TEAM;
    const int tib = threadIdx.x / TEAM;
    const int lane = threadIdx.x % TEAM;
    const long long task = (long long)blockIdx.x * TPB + tib;
    if (task >= (long long)n_sn * nact) return;
    const int J = sn_list[task % n_sn];
    const int b = act[task / n_sn];
    const int c = T.sn_ncols[J], m = T.sn_nrows[J], r = c + m, first = T.sn_first[J];
    double* xl = smem + (size_t)tib * (size_t)rmax;
    double* x = P.x + (size_t)b * T.n;
    const double* y = P.y + (size_t)b * T.n;
    const double* U = P.U + (size_t)b * T.u_size + T.sn_u_off[J];
    const int* rows = T.row_idx + T.sn_row_off[J];
    for (int a = lane; a < m; a += TEAM) xl[c + a] = x[rows[a]];
    team_sync<TEAM>();
    for (int k = lane; k < c; k += TEAM) {
        double s = y[first + k];
        const double* Uk = U + (size_t)k * r + c;
        for (int a = 0; a < m; ++a) s -= Uk[a] * xl[c + a];
        xl[k] = s;
    }
    team_sync<TEAM>();
    for (int k = c - 1; k >= 0; --k) {
        const double xk = xl[k] / U[(size_t)k * r + k];
        for (int i = lane; i < k; i += TEAM) xl[i] -= U[(size_t)i * r + k] * xk;
        if (lane == 0) x[first + k] = xk;
        team_sync<TEAM>();
    }
}

// Per-instance NR bookkeeping after a linear solve (one block per instance).
__global__ __launch_bounds__(256) void k_update(DevTopo T, DevBatch P, const int* __restrict__ act,
                                                int max_iter, int nsteps) {
#pragma clang fp contract(off)
    __shared__ double red[8];
    __shared__ int redi[8];
    const int b = act[blockIdx.x];
    const int n = T.n;
    InstState st = P.st[b];
    if (st.singular) {
        // singular Jacobian: abort this GMIN step and the solve (D15 accounting)
        if (threadIdx.x == 0) {
            st.total += st.it + 1;
            st.ref += 2 * (st.it + 1);
            st.prev_kcl = 1e9;
            st.status = NR_STATUS_SINGULAR;
            st.done = 1;
            P.st[b] = st;
        }
        return;
    }
    double* v = P.v + (size_t)b * n;
    double* vn = P.vn + (size_t)b * n;
    const double* x = P.x + (size_t)b * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double val = x[T.iperm[i]];
        if (isnan(val) || isinf(val)) val = 0.0;       // sanitize_vector
        vn[i] = val;
    }
    __syncthreads();
    // pnjlim in SoA order, dependency levels preserve the sequential semantics
    const double* Isb = P.Is + (size_t)b * T.nD;
    for (int l = 0; l < T.n_pn_levels; ++l) {
        for (int k = T.pn_level_ptr[l] + threadIdx.x; k < T.pn_level_ptr[l + 1]; k += blockDim.x) {
            const int i = T.pn_diode[k];
            const double Vt = T.d_pn_vt[i];
            const double Is = Isb[i] * T.d_pn_isf[i];
            const int p = T.d_a[i] - 1, q = T.d_c[i] - 1;
            double vcrit = Vt * log(Vt / (sqrt(2.0) * Is));
            double vjo = (p >= 0 ? v[p] : 0.0) - (q >= 0 ? v[q] : 0.0);
            double vjn = (p >= 0 ? vn[p] : 0.0) - (q >= 0 ? vn[q] : 0.0);
            double vl = pnjlim(vjn, vjo, Vt, vcrit);
            double corr = vjn - vl;
            if (fabs(corr) > 1e-12) {
                if (p >= 0 && q >= 0) { vn[p] -= corr * 0.5; vn[q] += corr * 0.5; }
                else if (p >= 0) vn[p] -= corr;
                else if (q >= 0) vn[q] += corr;
            }
        }
        __syncthreads();
    }
    double kcl = block_kcl(T, P, b, vn, red);
    if (st.it > 0 && kcl >= st.prev_kcl) {
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) vn[i] = v[i] + 0.5 * (vn[i] - v[i]);
            __syncthreads();
            kcl = block_kcl(T, P, b, vn, red);
            if (!(kcl >= st.prev_kcl)) break;
        }
    }
    st.prev_kcl = kcl;
    const double tol = (st.step == nsteps - 1) ? 1e-12 : 1.0;
    int conv;
    if (kcl < 1e-9) {
        conv = 1;
    } else {
        int ok = 1;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            double a = fabs(vn[i]), o = fabs(v[i]);
            double lim = 1e-3 * ((a < o) ? o : a) + 1e-6;
            if (fabs(vn[i] - v[i]) > lim) ok = 0;
        }
        ok = block_and(ok, redi);
        conv = ok && !(kcl > tol);
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] = vn[i];
    if (threadIdx.x == 0) {
        if (conv) {
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
                    st.status = NR_STATUS_NOT_CONVERGED; st.done = 1;
                } else { st.step++; st.it = 0; }
            }
        }
        P.st[b] = st;
    }
}

// rebuild the active list (single block)
__global__ void k_compact(DevBatch P) {
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
    for (int i = lo; i < hi; ++i) if (!P.st[i].done) P.act[pos++] = i;
    if (tid == nt - 1) *P.nact = cnt[tid];
}

__global__ __launch_bounds__(256) void k_final_kcl(DevTopo T, DevBatch P) {
    __shared__ double red[8];
    const int b = blockIdx.x;
    double k = block_kcl(T, P, b, P.v + (size_t)b * T.n, red);
    if (threadIdx.x == 0) P.st[b].kcl = k;
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
    mos_eval(volt(v, nv, g[i]), volt(v, nv, d[i]), volt(v, nv, s[i]), W[i], L[i], Kp[i], Vth[i], lam[i],
             pm[i], vgs[i], vds[i], ids[i], gm[i], gds[i]);
}

// ---------------------------------------------------------------------------
// host side: topology upload, batch, solve loop
// ---------------------------------------------------------------------------
namespace {

struct DevAlloc {
    std::vector<void*> ptrs;
    ~DevAlloc() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

template <typename Tv>
hipError_t upload(DevAlloc& al, const std::vector<Tv>& h, const Tv** out) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(1, h.size()) * sizeof(Tv);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return e;
    al.ptrs.push_back(p);
    if (!h.empty()) e = hipMemcpy(p, h.data(), h.size() * sizeof(Tv), hipMemcpyHostToDevice);
    *out = reinterpret_cast<const Tv*>(p);
    return e;
}

// fronts up to this order are factored by one wavefront (4 per block);
// larger ones by a 256-thread block
constexpr int SMALL_FRONT = 16;

struct LevelLists {
    std::vector<int> small_ptr, large_ptr;   // per level offsets
    std::vector<int> small_rmax, large_rmax;
    const int* d_small = nullptr;
    const int* d_large = nullptr;
};

}  // namespace

struct nr_topology {
    Topology t;
    nr_ctx* dev_ctx = nullptr;
    std::unique_ptr<DevAlloc> dev;
    DevTopo dt{};
    LevelLists lv;
};

static int ensure_device_topology(nr_ctx* ctx, nr_topology* tp) {
    if (tp->dev && tp->dev_ctx == ctx) return NR_OK;
    tp->dev.reset(new DevAlloc());
    DevAlloc& al = *tp->dev;
    const Topology& t = tp->t;
    DevTopo& d = tp->dt;
    d.n = t.n;
    d.nnz_a = t.nnz_a;
    d.nD = (int)t.soa.d_a.size();
    d.nM = (int)t.soa.m_d.size();
    d.nR = t.n_res;
    d.nV = t.n_vsrc;
    d.nI = t.n_isrc;
#define UP(field, vec) HIP_TRY(upload(al, vec, &d.field), ctx)
    UP(d_a, t.soa.d_a); UP(d_c, t.soa.d_c); UP(d_n, t.soa.d_n); UP(d_vt, t.soa.d_vt);
    UP(d_vtraw, t.soa.d_vtraw); UP(d_pn_vt, t.soa.d_pn_vt); UP(d_pn_isf, t.soa.d_pn_isf);
    UP(m_d, t.soa.m_d); UP(m_g, t.soa.m_g); UP(m_s, t.soa.m_s); UP(m_b, t.soa.m_b);
    UP(m_pmos, t.soa.m_pmos); UP(m_w, t.soa.m_w); UP(m_l, t.soa.m_l); UP(m_lam, t.soa.m_lam);
    UP(res_n1, t.res_n1); UP(res_n2, t.res_n2); UP(vsrc_pos, t.vsrc_pos); UP(vsrc_neg, t.vsrc_neg);
    UP(isrc_pos, t.isrc_pos); UP(isrc_neg, t.isrc_neg); UP(dio_anode_nl, t.dio_anode_nl);
    UP(lin_ptr, t.lin_ptr); UP(lin_term, t.lin_term); UP(nl_ptr, t.nl_ptr); UP(nl_term, t.nl_term);
    UP(slot_diag, t.slot_is_diag);
    UP(rhs_lin_ptr, t.rhs_lin_ptr); UP(rhs_lin_term, t.rhs_lin_term);
    UP(rhs_nl_ptr, t.rhs_nl_ptr); UP(rhs_nl_term, t.rhs_nl_term);
    UP(kcl_ptr, t.kcl_ptr); UP(kcl_term, t.kcl_term);
    UP(pn_level_ptr, t.pn_level_ptr); UP(pn_diode, t.pn_diode);
    d.n_pn_levels = (int)t.pn_level_ptr.size() - 1;
    UP(perm, t.perm); UP(iperm, t.iperm);
    {
        std::vector<int> f, c, m, ro, ab, ae, cb, ce;
        std::vector<long long> uo, co, ro2;
        for (const auto& s : t.sn) {
            f.push_back(s.first); c.push_back(s.ncols); m.push_back(s.nrows); ro.push_back(s.row_off);
            ab.push_back(s.a_begin); ae.push_back(s.a_end);
            cb.push_back(s.child_begin); ce.push_back(s.child_end);
            uo.push_back(s.u_off); co.push_back(s.cb_off); ro2.push_back(s.rcb_off);
        }
        UP(sn_first, f); UP(sn_ncols, c); UP(sn_nrows, m); UP(sn_row_off, ro);
        UP(sn_a_begin, ab); UP(sn_a_end, ae); UP(sn_child_begin, cb); UP(sn_child_end, ce);
        UP(sn_u_off, uo); UP(sn_cb_off, co); UP(sn_rcb_off, ro2);
    }
    UP(children, t.children); UP(row_idx, t.row_idx); UP(relmap, t.relmap);
    UP(fa_slot, t.front_a_slot); UP(fa_li, t.front_a_li); UP(fa_lj, t.front_a_lj);
#undef UP
    d.u_size = t.u_size;
    d.cb_size = t.cb_size;
    d.rcb_size = t.rcb_size;
    LevelLists& L = tp->lv;
    L = LevelLists();
    std::vector<int> small, large;
    L.small_ptr.assign(1, 0);
    L.large_ptr.assign(1, 0);
    for (int l = 0; l < t.n_levels; ++l) {
        int rs = 1, rl = 1;
        for (int k = t.level_ptr[l]; k < t.level_ptr[l + 1]; ++k) {
            const int J = t.level_sn[k];
            const int r = t.sn[J].ncols + t.sn[J].nrows;
            if (r <= SMALL_FRONT) { small.push_back(J); rs = std::max(rs, r); }
            else { large.push_back(J); rl = std::max(rl, r); }
        }
        L.small_ptr.push_back((int)small.size());
        L.large_ptr.push_back((int)large.size());
        L.small_rmax.push_back(rs);
        L.large_rmax.push_back(rl);
    }
    HIP_TRY(upload(al, small, &L.d_small), ctx);
    HIP_TRY(upload(al, large, &L.d_large), ctx);
    tp->dev_ctx = ctx;
    return NR_OK;
}

struct nr_batch {
    nr_ctx* ctx = nullptr;
    nr_topology* topo = nullptr;
    int B = 0;
    DevAlloc al;
    DevBatch p{};
    double* d_gmin = nullptr;
    bool has_seed = false;
    nr_timing timing{};
    std::vector<hipEvent_t> ev;
};

// largest front the LDS budget allows for the 256-thread factor kernel
static constexpr int MAX_LDS_FRONT = 140;   // 140*141*8 B = 158 KB

static int launch_factor_level(nr_batch* bt, int l, int nact, hipStream_t s) {
    nr_topology* tp = bt->topo;
    const LevelLists& L = tp->lv;
    const int ns = L.small_ptr[l + 1] - L.small_ptr[l];
    const int nl = L.large_ptr[l + 1] - L.large_ptr[l];
    if (ns > 0) {
        const int rmax = L.small_rmax[l];
        const long long tasks = (long long)ns * nact;
        const long long blocks = (tasks + 3) / 4;
        const size_t lds = (size_t)4 * (rmax * rmax + rmax) * sizeof(double);
        hipLaunchKernelGGL(k_factor<64>, dim3((unsigned)blocks), dim3(256), lds, s, tp->dt, bt->p, bt->p.act,
                           nact, L.d_small + L.small_ptr[l], ns, rmax);
    }
    if (nl > 0) {
        const int rmax = L.large_rmax[l];
        const long long tasks = (long long)nl * nact;
        const size_t lds = (size_t)(rmax * rmax + rmax) * sizeof(double);
        hipLaunchKernelGGL(k_factor<256>, dim3((unsigned)tasks), dim3(256), lds, s, tp->dt, bt->p, bt->p.act,
                           nact, L.d_large + L.large_ptr[l], nl, rmax);
    }
    return NR_OK;
}

static int launch_backsolve_level(nr_batch* bt, int l, int nact, hipStream_t s) {
    nr_topology* tp = bt->topo;
    const LevelLists& L = tp->lv;
    const int ns = L.small_ptr[l + 1] - L.small_ptr[l];
    const int nl = L.large_ptr[l + 1] - L.large_ptr[l];
    if (ns > 0) {
        const int rmax = L.small_rmax[l];
        const long long blocks = ((long long)ns * nact + 3) / 4;
        hipLaunchKernelGGL(k_backsolve<64>, dim3((unsigned)blocks), dim3(256), (size_t)4 * rmax * sizeof(double), s,
                           tp->dt, bt->p, bt->p.act, nact, L.d_small + L.small_ptr[l], ns, rmax);
    }
    if (nl > 0) {
        const int rmax = L.large_rmax[l];
        hipLaunchKernelGGL(k_backsolve<256>, dim3((unsigned)((long long)nl * nact)), dim3(256),
                           (size_t)rmax * sizeof(double), s, tp->dt, bt->p, bt->p.act, nact,
                           L.d_large + L.large_ptr[l], nl, rmax);
    }
    return NR_OK;
}

extern "C" {

NR_API const char* nr_version(void) { return "karam_nr 0.1 (gfx950)"; }

NR_API const char* nr_last_error(const nr_ctx* ctx) {
    if (ctx && !ctx->err.empty()) return ctx->err.c_str();
    return g_lib_error.c_str();
}

NR_API int nr_ctx_create(const nr_device_cfg* cfg, nr_ctx** out) {
    if (!out) return set_err(nullptr, NR_ERR_INVALID, "nr_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return set_err(nullptr, NR_ERR_NO_DEVICE, "no HIP device visible");
    std::unique_ptr<nr_ctx> c(new nr_ctx());
    c->device = cfg ? cfg->device : 0;
    c->verbose = cfg ? cfg->verbose : 0;
    if (c->device < 0 || c->device >= count) return set_err(nullptr, NR_ERR_INVALID, "device ordinal out of range");
    HIP_TRY(hipSetDevice(c->device), c.get());
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), c.get());
    *out = c.release();
    return NR_OK;
}

NR_API int nr_ctx_destroy(nr_ctx* ctx) {
    if (!ctx) return NR_OK;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NR_OK;
}

NR_API int nr_topology_build(nr_ctx* ctx, const nr_netlist* nl, nr_topology** out) {
    if (!nl || !out) return set_err(ctx, NR_ERR_INVALID, "nr_topology_build: NULL argument");
    *out = nullptr;
    std::unique_ptr<nr_topology> tp(new nr_topology());
    std::string err;
    int rc;
    try {
        rc = karam::build_topology(nl, tp->t, err);
    } catch (const std::exception& ex) {
        return set_err(ctx, NR_ERR_TOPOLOGY, std::string("topology: ") + ex.what());
    }
    if (rc != NR_OK) return set_err(ctx, rc, err);
    if (tp->t.max_front > MAX_LDS_FRONT)
        return set_err(ctx, NR_ERR_UNSUPPORTED, "frontal matrix larger than the LDS-resident limit");
    *out = tp.release();
    return NR_OK;
}

NR_API int nr_topology_destroy(nr_topology* topo) {
    delete topo;
    return NR_OK;
}

NR_API int nr_topology_get_info(const nr_topology* topo, nr_topology_info* info) {
    if (!topo || !info) return set_err(nullptr, NR_ERR_INVALID, "nr_topology_get_info: NULL argument");
    const Topology& t = topo->t;
    std::memset(info, 0, sizeof(*info));
    info->n = t.n;
    info->nnz_a = t.nnz_a;
    info->nnz_f = t.nnz_f;
    info->n_stamps = (int32_t)t.stamps.size();
    info->treewidth = t.treewidth;
    info->ordering_len = (int32_t)t.ordering_raw.size();
    info->fill_in = t.fill_in;
    info->n_supernodes = (int32_t)t.sn.size();
    info->n_levels = t.n_levels;
    info->max_front = t.max_front;
    info->n_res = t.n_res;
    info->n_vsrc = t.n_vsrc;
    info->n_isrc = t.n_isrc;
    info->n_cap = t.n_cap;
    info->n_diode = (int32_t)t.soa.d_a.size();
    info->n_mos = (int32_t)t.soa.m_d.size();
    info->lu_flops = t.lu_flops;
    info->u_entries = t.u_size;
    info->analyse_ms = t.analyse_ms;
    return NR_OK;
}

NR_API int nr_topology_export(const nr_topology* topo, int32_t what, int32_t* out, int64_t capacity,
                              int64_t* written) {
    if (!topo) return set_err(nullptr, NR_ERR_INVALID, "nr_topology_export: NULL topology");
    const Topology& t = topo->t;
    const std::vector<int32_t>* v = nullptr;
    switch (what) {
    case NR_EXPORT_ROW_PTR: v = &t.row_ptr; break;
    case NR_EXPORT_COL_IDX: v = &t.col_idx; break;
    case NR_EXPORT_STAMP_SLOT: v = &t.stamp_slot; break;
    case NR_EXPORT_ORDERING_RAW: v = &t.ordering_raw; break;
    case NR_EXPORT_PERM: v = &t.perm; break;
    case NR_EXPORT_FILL_ROW_PTR: v = &t.frow_ptr; break;
    case NR_EXPORT_FILL_COL_IDX: v = &t.fcol_idx; break;
    case NR_EXPORT_DIODE_ORDER: v = &t.soa.d_src; break;
    case NR_EXPORT_MOS_ORDER: v = &t.soa.m_src; break;
    default: return set_err(nullptr, NR_ERR_INVALID, "nr_topology_export: unknown item");
    }
    if (written) *written = (int64_t)v->size();
    if (!out) return NR_OK;
    if ((int64_t)v->size() > capacity)
        return set_err(nullptr, NR_ERR_INVALID, "nr_topology_export: capacity too small");
    if (!v->empty()) std::memcpy(out, v->data(), v->size() * sizeof(int32_t));
    return NR_OK;
}

NR_API void nr_solve_opts_default(nr_solve_opts* o) {
    if (!o) return;
    o->max_iter = 50;
    o->gmin_steps = 19;
    o->precision = NR_PREC_FP64;
    o->seeded_start_step = 18;
    o->record_timing = 0;
    o->use_graph = 0;
}

NR_API int nr_batch_alloc(nr_ctx* ctx, const nr_topology* topo_c, int32_t B, nr_batch** out) {
    if (!ctx || !topo_c || !out || B <= 0) return set_err(ctx, NR_ERR_INVALID, "nr_batch_alloc: bad argument");
    *out = nullptr;
    nr_topology* topo = const_cast<nr_topology*>(topo_c);
    HIP_TRY(hipSetDevice(ctx->device), ctx);
    int rc = ensure_device_topology(ctx, topo);
    if (rc) return rc;
    std::unique_ptr<nr_batch> bt(new nr_batch());
    bt->ctx = ctx;
    bt->topo = topo;
    bt->B = B;
    const Topology& t = topo->t;
    const size_t Bz = (size_t)B;
    const size_t nD = t.soa.d_a.size(), nM = t.soa.m_d.size();
    auto alloc = [&](double** p, size_t count) -> hipError_t {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(1, count) * sizeof(double));
        if (e == hipSuccess) {
            bt->al.ptrs.push_back(q);
            *p = (double*)q;
        }
        return e;
    };
    DevBatch& p = bt->p;
    p.B = B;
    HIP_TRY(alloc(&p.R, Bz * t.n_res), ctx);
    HIP_TRY(alloc(&p.Is, Bz * nD), ctx);
    HIP_TRY(alloc(&p.kp, Bz * nM), ctx);
    HIP_TRY(alloc(&p.vth, Bz * nM), ctx);
    HIP_TRY(alloc(&p.isrc, Bz * t.n_isrc), ctx);
    HIP_TRY(alloc(&p.vsrc, Bz * t.n_vsrc), ctx);
    HIP_TRY(alloc(&p.v, Bz * t.n), ctx);
    HIP_TRY(alloc(&p.vn, Bz * t.n), ctx);
    HIP_TRY(alloc(&p.glin, Bz * t.nnz_a), ctx);
    HIP_TRY(alloc(&p.rhslin, Bz * t.n), ctx);
    HIP_TRY(alloc(&p.A, Bz * t.nnz_a), ctx);
    HIP_TRY(alloc(&p.rhs, Bz * t.n), ctx);
    HIP_TRY(alloc(&p.dg, Bz * nD), ctx);
    HIP_TRY(alloc(&p.dieq, Bz * nD), ctx);
    HIP_TRY(alloc(&p.mgds, Bz * nM), ctx);
    HIP_TRY(alloc(&p.mgm, Bz * nM), ctx);
    HIP_TRY(alloc(&p.mieq, Bz * nM), ctx);
    HIP_TRY(alloc(&p.U, Bz * (size_t)t.u_size), ctx);
    HIP_TRY(alloc(&p.y, Bz * t.n), ctx);
    HIP_TRY(alloc(&p.x, Bz * t.n), ctx);
    HIP_TRY(alloc(&p.CB, Bz * (size_t)t.cb_size), ctx);
    HIP_TRY(alloc(&p.RCB, Bz * (size_t)t.rcb_size), ctx);
    HIP_TRY(alloc(&bt->d_gmin, 64), ctx);
    p.seed = nullptr;
    {
        void* q = nullptr;
        HIP_TRY(hipMalloc(&q, Bz * sizeof(InstState)), ctx);
        bt->al.ptrs.push_back(q);
        p.st = (InstState*)q;
        HIP_TRY(hipMalloc(&q, Bz * sizeof(int)), ctx);
        bt->al.ptrs.push_back(q);
        p.act = (int*)q;
        HIP_TRY(hipMalloc(&q, sizeof(int)), ctx);
        bt->al.ptrs.push_back(q);
        p.nact = (int*)q;
    }
    // every instance starts from the template netlist's parameters (SoA order
    // for Is / KP / VTO, netlist order for the rest)
    {
        std::vector<double> h;
        auto fill = [&](double* dst, const std::vector<double>& one) -> hipError_t {
            if (one.empty()) return hipSuccess;
            h.resize(one.size() * Bz);
            for (size_t b = 0; b < Bz; ++b) std::copy(one.begin(), one.end(), h.begin() + b * one.size());
            return hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
        };
        HIP_TRY(fill(p.R, t.res_ohms), ctx);
        HIP_TRY(fill(p.Is, t.soa.d_is), ctx);
        HIP_TRY(fill(p.kp, t.soa.m_kp), ctx);
        HIP_TRY(fill(p.vth, t.soa.m_vth), ctx);
        HIP_TRY(fill(p.isrc, t.isrc_amps), ctx);
        HIP_TRY(fill(p.vsrc, t.vsrc_volts), ctx);
    }
    *out = bt.release();
    return NR_OK;
}

NR_API int nr_batch_free(nr_batch* batch) {
    if (!batch) return NR_OK;
    (void)hipSetDevice(batch->ctx->device);
    for (hipEvent_t e : batch->ev) (void)hipEventDestroy(e);
    delete batch;
    return NR_OK;
}

NR_API int nr_batch_set_params(nr_batch* bt, int32_t field, const double* values, int64_t count) {
    if (!bt || !values) return set_err(bt ? bt->ctx : nullptr, NR_ERR_INVALID, "nr_batch_set_params: NULL argument");
    nr_ctx* ctx = bt->ctx;
    const Topology& t = bt->topo->t;
    const size_t Bz = (size_t)bt->B;
    double* dst = nullptr;
    size_t per = 0;
    const std::vector<int32_t>* soa_src = nullptr;   // netlist -> SoA permutation
    switch (field) {
    case NR_FIELD_RES_OHMS: dst = bt->p.R; per = t.n_res; break;
    case NR_FIELD_DIO_IS: dst = bt->p.Is; per = t.soa.d_a.size(); soa_src = &t.soa.d_src; break;
    case NR_FIELD_MOS_KP: dst = bt->p.kp; per = t.soa.m_d.size(); soa_src = &t.soa.m_src; break;
    case NR_FIELD_MOS_VTO: dst = bt->p.vth; per = t.soa.m_d.size(); soa_src = &t.soa.m_src; break;
    case NR_FIELD_ISRC_AMPS: dst = bt->p.isrc; per = t.n_isrc; break;
    case NR_FIELD_VSRC_VOLTS: dst = bt->p.vsrc; per = t.n_vsrc; break;
    case NR_FIELD_CAP_FARADS: return per == 0 ? NR_OK : NR_OK;   // DC: capacitors are open
    default: return set_err(ctx, NR_ERR_INVALID, "nr_batch_set_params: unknown field");
    }
    if ((size_t)count != per * Bz) return set_err(ctx, NR_ERR_INVALID, "nr_batch_set_params: count != B * field size");
    if (per == 0) return NR_OK;
    HIP_TRY(hipSetDevice(ctx->device), ctx);
    if (soa_src) {
        std::vector<double> h(per * Bz);
        for (size_t b = 0; b < Bz; ++b)
            for (size_t pidx = 0; pidx < per; ++pidx) h[b * per + pidx] = values[b * per + (*soa_src)[pidx]];
        HIP_TRY(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), ctx);
    } else {
        HIP_TRY(hipMemcpy(dst, values, per * Bz * sizeof(double), hipMemcpyHostToDevice), ctx);
    }
    return NR_OK;
}

NR_API int nr_batch_set_seed(nr_batch* bt, const double* v0, int64_t count) {
    if (!bt || !v0) return set_err(bt ? bt->ctx : nullptr, NR_ERR_INVALID, "nr_batch_set_seed: NULL argument");
    nr_ctx* ctx = bt->ctx;
    const size_t need = (size_t)bt->B * bt->topo->t.n;
    if ((size_t)count != need) return set_err(ctx, NR_ERR_INVALID, "nr_batch_set_seed: count != B * n");
    HIP_TRY(hipSetDevice(ctx->device), ctx);
    if (!bt->has_seed) {
        void* q = nullptr;
        HIP_TRY(hipMalloc(&q, need * sizeof(double)), ctx);
        bt->al.ptrs.push_back(q);
        bt->p.seed = (double*)q;
        bt->has_seed = true;
    }
    HIP_TRY(hipMemcpy(bt->p.seed, v0, need * sizeof(double), hipMemcpyHostToDevice), ctx);
    return NR_OK;
}

NR_API int nr_batch_clear_seed(nr_batch* bt) {
    if (!bt) return set_err(nullptr, NR_ERR_INVALID, "nr_batch_clear_seed: NULL batch");
    bt->p.seed = nullptr;   // buffer stays allocated for a later seed
    bt->has_seed = false;
    return NR_OK;
}

NR_API int nr_batch_solve_dc(nr_batch* bt, const nr_solve_opts* opts_in, nr_result* res) {
    if (!bt) return set_err(nullptr, NR_ERR_INVALID, "nr_batch_solve_dc: NULL batch");
    nr_ctx* ctx = bt->ctx;
    nr_solve_opts o;
    if (opts_in) o = *opts_in;
    else nr_solve_opts_default(&o);
    if (o.max_iter <= 0 || o.gmin_steps <= 0 || o.gmin_steps > 60)
        return set_err(ctx, NR_ERR_INVALID, "nr_batch_solve_dc: bad max_iter / gmin_steps");
    if (o.precision != NR_PREC_FP64)
        return set_err(ctx, NR_ERR_UNSUPPORTED, "nr_batch_solve_dc: only FP64 precision is built");
    if (o.seeded_start_step < 0 || o.seeded_start_step >= o.gmin_steps)
        return set_err(ctx, NR_ERR_INVALID, "nr_batch_solve_dc: seeded_start_step out of range");
    HIP_TRY(hipSetDevice(ctx->device), ctx);
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = ctx->stream;
    nr_topology* tp = bt->topo;
    const Topology& t = tp->t;
    // GMIN schedule of solveDC (solvers/circuitsim_solver.cpp:235-240)
    std::vector<double> gtab(64, 1e-12);
    for (int k = 0; k < o.gmin_steps; ++k) {
        double g = std::pow(10.0, -(3.0 + k * 0.5));
        if (g < 1e-12) g = 1e-12;
        gtab[k] = g;
    }
    HIP_TRY(hipMemcpyAsync(bt->d_gmin, gtab.data(), 64 * sizeof(double), hipMemcpyHostToDevice, s), ctx);
    if (!bt->has_seed) bt->p.seed = nullptr;
    hipLaunchKernelGGL(k_init, dim3(bt->B), dim3(256), 0, s, tp->dt, bt->p, o.seeded_start_step);
    HIP_TRY(hipGetLastError(), ctx);
    std::memset(&bt->timing, 0, sizeof(bt->timing));
    if (o.record_timing && bt->ev.empty()) {
        bt->ev.resize(5);
        for (auto& e : bt->ev) HIP_TRY(hipEventCreate(&e), ctx);
    }
    int nact = bt->B;
    int sweeps = 0;
    double evals = 0.0;
    const int max_sweeps = o.gmin_steps * (o.max_iter + 1) + 4;
    const int nD = (int)t.soa.d_a.size(), nM = (int)t.soa.m_d.size();
    while (nact > 0 && sweeps < max_sweeps) {
        if (o.record_timing) HIP_TRY(hipEventRecord(bt->ev[0], s), ctx);
        {
            const int nd = nD + nM;
            if (nd > 0)
                hipLaunchKernelGGL(k_eval, dim3((nd + 255) / 256, nact), dim3(256), 0, s, tp->dt, bt->p, bt->p.act);
            const int na = std::max(t.nnz_a, t.n);
            hipLaunchKernelGGL(k_assemble, dim3((na + 255) / 256, nact), dim3(256), 0, s, tp->dt, bt->p, bt->p.act,
                               bt->d_gmin);
        }
        if (o.record_timing) HIP_TRY(hipEventRecord(bt->ev[1], s), ctx);
        for (int l = 0; l < t.n_levels; ++l) launch_factor_level(bt, l, nact, s);
        if (o.record_timing) HIP_TRY(hipEventRecord(bt->ev[2], s), ctx);
        for (int l = t.n_levels - 1; l >= 0; --l) launch_backsolve_level(bt, l, nact, s);
        if (o.record_timing) HIP_TRY(hipEventRecord(bt->ev[3], s), ctx);
        hipLaunchKernelGGL(k_update, dim3(nact), dim3(256), 0, s, tp->dt, bt->p, bt->p.act, o.max_iter, o.gmin_steps);
        hipLaunchKernelGGL(k_compact, dim3(1), dim3(1024), 0, s, bt->p);
        if (o.record_timing) HIP_TRY(hipEventRecord(bt->ev[4], s), ctx);
        HIP_TRY(hipGetLastError(), ctx);
        int next = 0;
        HIP_TRY(hipMemcpyAsync(&next, bt->p.nact, sizeof(int), hipMemcpyDeviceToHost, s), ctx);
        HIP_TRY(hipStreamSynchronize(s), ctx);
        if (o.record_timing) {
            float ms;
            for (int ph = 0; ph < NR_PHASE_COUNT; ++ph) {
                HIP_TRY(hipEventElapsedTime(&ms, bt->ev[ph], bt->ev[ph + 1]), ctx);
                bt->timing.ms[ph] += ms;
                bt->timing.launches[ph] += 1;
            }
            bt->timing.factor_instances += nact;
        }
        evals += (double)nact * (nD + nM);
        nact = next;
        ++sweeps;
    }
    hipLaunchKernelGGL(k_final_kcl, dim3(bt->B), dim3(256), 0, s, tp->dt, bt->p);
    HIP_TRY(hipGetLastError(), ctx);
    HIP_TRY(hipStreamSynchronize(s), ctx);
    auto t1 = std::chrono::steady_clock::now();
    if (res) {
        std::vector<InstState> hs(bt->B);
        HIP_TRY(hipMemcpy(hs.data(), bt->p.st, hs.size() * sizeof(InstState), hipMemcpyDeviceToHost), ctx);
        std::memset(res, 0, sizeof(*res));
        res->B = bt->B;
        double sum = 0.0;
        for (const auto& x : hs) {
            if (x.status == NR_STATUS_CONVERGED) {
                res->n_converged++;
                res->max_kcl = std::max(res->max_kcl, x.kcl);
            }
            res->max_iterations = std::max(res->max_iterations, x.total);
            sum += x.total;
        }
        res->mean_iterations = sum / bt->B;
        res->batched_iterations = sweeps;
        res->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        res->device_evals = evals;
    }
    bt->timing.factor_kernel_ms = bt->timing.ms[NR_PHASE_FACTOR];
    return NR_OK;
}

NR_API int nr_batch_get_voltages(const nr_batch* bt, double* out, int64_t count) {
    if (!bt || !out) return set_err(nullptr, NR_ERR_INVALID, "nr_batch_get_voltages: NULL argument");
    const size_t need = (size_t)bt->B * bt->topo->t.n;
    if ((size_t)count != need) return set_err(bt->ctx, NR_ERR_INVALID, "nr_batch_get_voltages: count != B * n");
    HIP_TRY(hipSetDevice(bt->ctx->device), bt->ctx);
    HIP_TRY(hipMemcpy(out, bt->p.v, need * sizeof(double), hipMemcpyDeviceToHost), bt->ctx);
    return NR_OK;
}

NR_API int nr_batch_get_stats(const nr_batch* bt, nr_instance_stats* out, int64_t count) {
    if (!bt || !out) return set_err(nullptr, NR_ERR_INVALID, "nr_batch_get_stats: NULL argument");
    if (count != bt->B) return set_err(bt->ctx, NR_ERR_INVALID, "nr_batch_get_stats: count != B");
    HIP_TRY(hipSetDevice(bt->ctx->device), bt->ctx);
    std::vector<InstState> hs(bt->B);
    HIP_TRY(hipMemcpy(hs.data(), bt->p.st, hs.size() * sizeof(InstState), hipMemcpyDeviceToHost), bt->ctx);
    for (int b = 0; b < bt->B; ++b) {
        out[b].iterations = hs[b].total;
        out[b].ref_iterations = hs[b].ref;
        out[b].status = hs[b].status;
        out[b].gmin_steps_done = hs[b].step + (hs[b].status == NR_STATUS_CONVERGED ? 1 : 0);
        out[b].kcl = hs[b].kcl;
    }
    return NR_OK;
}

NR_API int nr_batch_get_timing(const nr_batch* bt, nr_timing* out) {
    if (!bt || !out) return set_err(nullptr, NR_ERR_INVALID, "nr_batch_get_timing: NULL argument");
    *out = bt->timing;
    return NR_OK;
}

NR_API int nr_eval_diodes(nr_ctx* ctx, int32_t count, const int32_t* node_a, const int32_t* node_c,
                          const double* Is, const double* N, const double* Vt, const double* voltages,
                          int32_t n_voltages, double* v_d, double* i_d, double* g_d) {
    if (!ctx) return set_err(nullptr, NR_ERR_INVALID, "nr_eval_diodes: NULL context");
    if (count < 0 || n_voltages < 0) return set_err(ctx, NR_ERR_INVALID, "nr_eval_diodes: negative size");
    if (count == 0) return NR_OK;
    if (!node_a || !node_c || !Is || !N || !Vt || (!voltages && n_voltages) || !v_d || !i_d || !g_d)
        return set_err(ctx, NR_ERR_INVALID, "nr_eval_diodes: NULL array");
    HIP_TRY(hipSetDevice(ctx->device), ctx);
    DevAlloc al;
    auto dup = [&](const void* h, size_t bytes, void** d) -> hipError_t {
        hipError_t e = hipMalloc(d, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) return e;
        al.ptrs.push_back(*d);
        return h && bytes ? hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    void *da, *dc, *dis, *dn, *dvt, *dv, *o1, *o2, *o3;
    const size_t c4 = (size_t)count * 4, c8 = (size_t)count * 8;
    HIP_TRY(dup(node_a, c4, &da), ctx); HIP_TRY(dup(node_c, c4, &dc), ctx);
    HIP_TRY(dup(Is, c8, &dis), ctx); HIP_TRY(dup(N, c8, &dn), ctx); HIP_TRY(dup(Vt, c8, &dvt), ctx);
    HIP_TRY(dup(voltages, (size_t)n_voltages * 8, &dv), ctx);
    HIP_TRY(dup(nullptr, c8, &o1), ctx); HIP_TRY(dup(nullptr, c8, &o2), ctx); HIP_TRY(dup(nullptr, c8, &o3), ctx);
    hipLaunchKernelGGL(k_eval_diodes_slot, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, count,
                       (const int*)da, (const int*)dc, (const double*)dis, (const double*)dn, (const double*)dvt,
                       (const double*)dv, n_voltages, (double*)o1, (double*)o2, (double*)o3);
    HIP_TRY(hipGetLastError(), ctx);
    HIP_TRY(hipStreamSynchronize(ctx->stream), ctx);
    HIP_TRY(hipMemcpy(v_d, o1, c8, hipMemcpyDeviceToHost), ctx);
    HIP_TRY(hipMemcpy(i_d, o2, c8, hipMemcpyDeviceToHost), ctx);
    HIP_TRY(hipMemcpy(g_d, o3, c8, hipMemcpyDeviceToHost), ctx);
    return NR_OK;
}

NR_API int nr_eval_mosfets(nr_ctx* ctx, int32_t count, const int32_t* drains, const int32_t* gates,
                           const int32_t* sources, const double* W, const double* L, const double* Kp,
                           const double* Vth, const double* lambda, const int32_t* isPMOS,
                           const double* voltages, int32_t n_voltages, double* vgs, double* vds, double* ids,
                           double* gm, double* gds) {
    if (!ctx) return set_err(nullptr, NR_ERR_INVALID, "nr_eval_mosfets: NULL context");
    if (count < 0 || n_voltages < 0) return set_err(ctx, NR_ERR_INVALID, "nr_eval_mosfets: negative size");
    if (count == 0) return NR_OK;
    if (!drains || !gates || !sources || !W || !L || !Kp || !Vth || !lambda || !isPMOS ||
        (!voltages && n_voltages) || !vgs || !vds || !ids || !gm || !gds)
        return set_err(ctx, NR_ERR_INVALID, "nr_eval_mosfets: NULL array");
    HIP_TRY(hipSetDevice(ctx->device), ctx);
    DevAlloc al;
    auto dup = [&](const void* h, size_t bytes, void** d) -> hipError_t {
        hipError_t e = hipMalloc(d, std::max<size_t>(bytes, 8));
        if (e != hipSuccess) return e;
        al.ptrs.push_back(*d);
        return h && bytes ? hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    const size_t c4 = (size_t)count * 4, c8 = (size_t)count * 8;
    void *dd, *dg, *ds, *dw, *dl, *dkp, *dvth, *dlam, *dpm, *dv, *o[5];
    HIP_TRY(dup(drains, c4, &dd), ctx); HIP_TRY(dup(gates, c4, &dg), ctx); HIP_TRY(dup(sources, c4, &ds), ctx);
    HIP_TRY(dup(W, c8, &dw), ctx); HIP_TRY(dup(L, c8, &dl), ctx); HIP_TRY(dup(Kp, c8, &dkp), ctx);
    HIP_TRY(dup(Vth, c8, &dvth), ctx); HIP_TRY(dup(lambda, c8, &dlam), ctx); HIP_TRY(dup(isPMOS, c4, &dpm), ctx);
    HIP_TRY(dup(voltages, (size_t)n_voltages * 8, &dv), ctx);
    for (int k = 0; k < 5; ++k) HIP_TRY(dup(nullptr, c8, &o[k]), ctx);
    hipLaunchKernelGGL(k_eval_mos_slot, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, count,
                       (const int*)dd, (const int*)dg, (const int*)ds, (const double*)dw, (const double*)dl,
                       (const double*)dkp, (const double*)dvth, (const double*)dlam, (const int*)dpm,
                       (const double*)dv, n_voltages, (double*)o[0], (double*)o[1], (double*)o[2], (double*)o[3],
                       (double*)o[4]);
    HIP_TRY(hipGetLastError(), ctx);
    HIP_TRY(hipStreamSynchronize(ctx->stream), ctx);
    double* outs[5] = {vgs, vds, ids, gm, gds};
    for (int k = 0; k < 5; ++k) HIP_TRY(hipMemcpy(outs[k], o[k], c8, hipMemcpyDeviceToHost), ctx);
    return NR_OK;
}

}  // extern "C"
