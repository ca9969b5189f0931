/*
 * oracle.c — CPU FP64 restatement of the reference's batched Newton-Raphson
 * path: the DC solve (GMIN stepping, pnjlim, damping, KCL audit, convergence
 * diagnostics), the transient (companions, fixed steps and the
 * IntegratorArbiter's step control), the AC sweep and the noise analysis.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Built without fast-math and with
 * FP contraction off.  Each function follows the reference file:line named in
 * its comment, and its comment states where it departs from the reference's
 * literal behaviour.  Every function is pinned by a test: golden vectors of
 * the reference's own translation units (oracle/_ref), closed forms or
 * brute-force checks; the pins are listed where each test file begins.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GMIN_DEV 1e-12

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */
static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}
static void* xmalloc(size_t sz) {
    void* p = malloc(sz ? sz : 1);
    if (!p) abort();
    return p;
}
static double dmin_ref(double a, double b) { return (b < a) ? b : a; }  /* std::min(a,b) */
static double dmax_ref(double a, double b) { return (a < b) ? b : a; }  /* std::max(a,b) */

/* ------------------------------------------------------------------------ */
/* a2: tensorizeNetlist (tensors/physics_tensors.cpp:305-493)               */
/* ------------------------------------------------------------------------ */
struct orc_soa {
    int32_t n;
    int32_t nD, nM;
    int32_t *da, *dc, *d_src;
    double *dIs, *dN, *dVt;
    int32_t *md, *mg, *ms, *mb, *mP, *m_src;
    double *mW, *mL, *mKp, *mVth, *mLam;
};

static const char* model_name(const nr_netlist* nl, int32_t idx, const char* dflt) {
    if (idx < 0 || idx >= nl->n_models || !nl->models[idx].name) return dflt;
    return nl->models[idx].name;
}

/* Find the card named `name` (first match), NULL if none. */
static const nr_model_card* find_card(const nr_netlist* nl, const char* name) {
    for (int32_t i = 0; i < nl->n_models; ++i) {
        const nr_model_card* c = &nl->models[i];
        if (c->type && c->name && strcmp(c->name, name) == 0) return c;
    }
    return NULL;
}
static int has_any_card(const nr_netlist* nl) {
    for (int32_t i = 0; i < nl->n_models; ++i)
        if (nl->models[i].type) return 1;
    return 0;
}

/* stable sort of idx[] by name via insertion/merge (std::map bucket order:
 * lexicographic model name, netlist order within a bucket). */
typedef struct { const char* key; int32_t idx; } name_item;
static void stable_sort_names(name_item* it, int32_t n) {
    if (n < 2) return;
    name_item* tmp = (name_item*)xmalloc(sizeof(name_item) * (size_t)n);
    for (int32_t w = 1; w < n; w *= 2) {
        for (int32_t lo = 0; lo < n; lo += 2 * w) {
            int32_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int32_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) {
                if (strcmp(it[j].key, it[i].key) < 0) tmp[k++] = it[j++];
                else tmp[k++] = it[i++];
            }
            while (i < mid) tmp[k++] = it[i++];
            while (j < hi) tmp[k++] = it[j++];
        }
        memcpy(it, tmp, sizeof(name_item) * (size_t)n);
    }
    free(tmp);
}

orc_soa* orc_tensorize(const nr_netlist* nl) {
    orc_soa* s = (orc_soa*)xcalloc(1, sizeof(orc_soa));
    s->n = nl->num_nodes;
    const double tscale = nl->ambient_temp_K / 300.15;   /* physics_tensors.cpp:419 */
    /* diodes: bucket by modelName (default "diode_generic", circuit.h:404) */
    s->nD = nl->n_diode;
    name_item* items = (name_item*)xmalloc(sizeof(name_item) * (size_t)(s->nD + 1));
    for (int32_t i = 0; i < s->nD; ++i) {
        items[i].key = model_name(nl, nl->dio_model ? nl->dio_model[i] : -1, "diode_generic");
        items[i].idx = i;
    }
    stable_sort_names(items, s->nD);
    s->da = (int32_t*)xcalloc((size_t)s->nD, 4); s->dc = (int32_t*)xcalloc((size_t)s->nD, 4);
    s->d_src = (int32_t*)xcalloc((size_t)s->nD, 4);
    s->dIs = (double*)xcalloc((size_t)s->nD, 8); s->dN = (double*)xcalloc((size_t)s->nD, 8);
    s->dVt = (double*)xcalloc((size_t)s->nD, 8);
    for (int32_t p = 0; p < s->nD; ++p) {
        int32_t i = items[p].idx;
        s->d_src[p] = i;
        s->da[p] = nl->dio_anode[i]; s->dc[p] = nl->dio_cathode[i];
        s->dIs[p] = nl->dio_is[i]; s->dN[p] = nl->dio_n[i];
        s->dVt[p] = nl->dio_vt[i] * tscale;
    }
    free(items);
    /* MOSFETs: bucket by modelName; parameters from the card (:424-441 via
     * collectFromBlock :407-420); no cards at all -> NMOS defaults. */
    s->nM = nl->n_mos;
    items = (name_item*)xmalloc(sizeof(name_item) * (size_t)(s->nM + 1));
    for (int32_t i = 0; i < s->nM; ++i) {
        items[i].key = model_name(nl, nl->mos_model ? nl->mos_model[i] : -1, "");
        items[i].idx = i;
    }
    stable_sort_names(items, s->nM);
    size_t m = (size_t)s->nM;
    s->md = (int32_t*)xcalloc(m, 4); s->mg = (int32_t*)xcalloc(m, 4); s->ms = (int32_t*)xcalloc(m, 4);
    s->mb = (int32_t*)xcalloc(m, 4); s->mP = (int32_t*)xcalloc(m, 4); s->m_src = (int32_t*)xcalloc(m, 4);
    s->mW = (double*)xcalloc(m, 8); s->mL = (double*)xcalloc(m, 8); s->mKp = (double*)xcalloc(m, 8);
    s->mVth = (double*)xcalloc(m, 8); s->mLam = (double*)xcalloc(m, 8);
    const int cards = has_any_card(nl);
    for (int32_t p = 0; p < s->nM; ++p) {
        int32_t i = items[p].idx;
        const char* name = items[p].key;
        int pm = 0;
        double kp = 200e-6, vth = 0.7, lam = 0.02;
        if (cards) {
            const nr_model_card* c = find_card(nl, name);
            if (c) {
                pm = strcmp(c->type, "PMOS") == 0;
                kp = c->has_kp ? c->kp : (pm ? 100e-6 : 200e-6);
                vth = c->has_vto ? c->vto : (pm ? -0.7 : 0.7);
                lam = c->has_lambda ? c->lambda : 0.02;
            } else {
                pm = (strstr(name, "PMOS") != NULL) || (strstr(name, "pmos") != NULL);
                if (pm) { kp = 100e-6; vth = -0.7; }
            }
        }
        /* per-device Monte-Carlo KP / VTO overrides (an extension) */
        if (nl->mos_kp) kp = nl->mos_kp[i];
        if (nl->mos_vto) vth = nl->mos_vto[i];
        s->m_src[p] = i;
        s->md[p] = nl->mos_d[i]; s->mg[p] = nl->mos_g[i]; s->ms[p] = nl->mos_s[i]; s->mb[p] = nl->mos_b[i];
        s->mW[p] = nl->mos_w[i]; s->mL[p] = nl->mos_l[i];
        s->mKp[p] = kp; s->mVth[p] = vth; s->mLam[p] = lam; s->mP[p] = pm;
    }
    free(items);
    return s;
}

void orc_soa_free(orc_soa* s) {
    if (!s) return;
    free(s->da); free(s->dc); free(s->d_src); free(s->dIs); free(s->dN); free(s->dVt);
    free(s->md); free(s->mg); free(s->ms); free(s->mb); free(s->mP); free(s->m_src);
    free(s->mW); free(s->mL); free(s->mKp); free(s->mVth); free(s->mLam);
    free(s);
}

/* ------------------------------------------------------------------------ */
/* a4: batchDiodePhysics_generic (tensors/physics_tensors.cpp:19-51)        */
/* ------------------------------------------------------------------------ */
static inline double volt(const double* v, int32_t nv, int32_t node) {
    return (node > 0 && node <= nv) ? v[node - 1] : 0.0;
}

void orc_eval_diodes(int32_t count, const int32_t* a, const int32_t* c,
                     const double* Is, const double* N, const double* Vt,
                     const double* v, int32_t nv,
                     double* vd, double* id, double* gd) {
    for (int32_t i = 0; i < count; ++i) {
        double v_d = volt(v, nv, a[i]) - volt(v, nv, c[i]);
        vd[i] = v_d;
        double n_vt = N[i] * Vt[i];
        double vcrit = 30.0 * n_vt;
        double arg = v_d / n_vt;
        double arg_c = dmin_ref(arg, 30.0);
        double e = exp(arg_c);
        double i_base = Is[i] * (e - 1.0);
        double g_base = (Is[i] / n_vt) * e;
        double delta = dmax_ref(0.0, v_d - vcrit);
        id[i] = i_base + delta * g_base;
        gd[i] = dmax_ref(GMIN_DEV, g_base);
    }
}

/* ------------------------------------------------------------------------ */
/* a5: batchMosfetPhysics_generic (tensors/physics_tensors.cpp:53-104)      */
/* ------------------------------------------------------------------------ */
static inline void mos_one(double vg, double vd_, double vs, double W, double L,
                           double Kp, double Vth, double lam, int pmos,
                           double* o_vgs, double* o_vds, double* o_ids,
                           double* o_gm, double* o_gds) {
    double sign = pmos ? -1.0 : 1.0;
    double vgs = vg - vs, vds = vd_ - vs;
    double lvgs = sign * vgs, lvds = sign * vds, lvth = sign * Vth;
    double K = 0.5 * Kp * (W / L);
    double vov = lvgs - lvth;
    double ids, gm, gds;
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
    *o_vgs = vgs; *o_vds = vds; *o_ids = ids; *o_gm = gm; *o_gds = gds;
}

void orc_eval_mosfets(int32_t count, const int32_t* d, const int32_t* g,
                      const int32_t* s, const double* W, const double* L,
                      const double* Kp, const double* Vth, const double* lam,
                      const int32_t* pmos, const double* v, int32_t nv,
                      double* vgs, double* vds, double* ids, double* gm,
                      double* gds) {
    for (int32_t i = 0; i < count; ++i)
        mos_one(volt(v, nv, g[i]), volt(v, nv, d[i]), volt(v, nv, s[i]), W[i], L[i],
                Kp[i], Vth[i], lam[i], pmos[i], &vgs[i], &vds[i], &ids[i], &gm[i], &gds[i]);
}

/* ------------------------------------------------------------------------ */
/* FP32 path (device evaluation and LU in single precision, FP64 defect    */
/* correction): the Jacobian entries from the generic-kernel equations      */
/* (tensors/physics_tensors.cpp:19-104) evaluated in float.  Junction       */
/* voltages are differenced in FP64 and rounded once; parameters are        */
/* rounded to float; the currents are not needed (the rhs is the FP64 KCL). */
/* ------------------------------------------------------------------------ */
static float diode_g_f32(double vd64, double Is, double N, double Vt) {
    const float vd = (float)vd64, n_vt = (float)N * (float)Vt;
    const float arg = vd / n_vt;
    const float argc = (30.0f < arg) ? 30.0f : arg;
    const float gb = ((float)Is / n_vt) * expf(argc);
    return (1e-12f < gb) ? gb : 1e-12f;
}

static void mos_g_f32(double vgs64, double vds64, double W, double L, double Kp, double Vth, double lam,
                      int pmos, float* o_gm, float* o_gds) {
    const float sign = pmos ? -1.0f : 1.0f;
    const float lvgs = sign * (float)vgs64, lvds = sign * (float)vds64, lvth = sign * (float)Vth;
    const float K = 0.5f * (float)Kp * ((float)W / (float)L);
    const float vov = lvgs - lvth;
    float gm, gds;
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
    *o_gm = gm; *o_gds = gds;
}

/* ------------------------------------------------------------------------ */
/* Topology: stamps, pattern, Min-Fill, symbolic                            */
/* ------------------------------------------------------------------------ */
struct orc_topo {
    int32_t n;
    orc_soa* soa;                  /* SoA of the template netlist */
    int32_t n_stamps;
    int32_t *st_row, *st_col;      /* triplet coordinates in emission order */
    int32_t nnz_a;
    int32_t *row_ptr, *col_idx;    /* CSR pattern (createCsr) */
    int32_t* stamp_slot;
    int32_t n_vars, ordering_len, treewidth;
    int32_t* ordering_raw;         /* Min-Fill ids incl. branch ids */
    int32_t *perm, *iperm;
    int32_t nnz_f, fill_in;
    int32_t *frow_ptr, *fcol_idx, *fdiag;
    /* relaxed supernodes of the fill pattern (multifrontal LU) */
    int32_t ns;
    int32_t *sn_first, *sn_ncols, *sn_nrows, *sn_row_off, *sn_rows;
    int32_t *sn_child_ptr, *sn_child;
    int32_t *sn_a_ptr, *sn_a_slot, *sn_a_li, *sn_a_lj;
    int32_t transient;             /* capacitor companions in the pattern */
};

/* growable int vector */
typedef struct { int32_t* a; int32_t n, cap; } ivec;
static void iv_push(ivec* v, int32_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 8;
        v->a = (int32_t*)realloc(v->a, sizeof(int32_t) * (size_t)v->cap);
        if (!v->a) abort();
    }
    v->a[v->n++] = x;
}

/* AC conductance of a diode (buildAdmittanceMatrix, solvers/ac_solver.h:
 * 186-196): the derivative of diode_current_clamped as its dual numbers
 * compute it (the dual quotient vd / n_vt carries n_vt / n_vt^2; past
 * arg = 30 the continuation's slope (Is / n_vt) e^30), at least 1e-15. */
static double ac_diode_g(double vd, double Is, double N, double Vt) {
    const double n_vt = N * Vt;
    const double arg = vd / n_vt;
    const double g = arg > 30.0 ? (Is / n_vt) * exp(30.0) : Is * (exp(arg) * (n_vt / (n_vt * n_vt)));
    return g < 1e-15 ? 1e-15 : g;
}

/* Emit the DC triplet coordinates in stampSoABlock order
 * (physics/circuitsim.h:2413-2751, h = 0) followed by the gmin diagonal
 * (addDiagonalConditioning, :2857-2860).  When vals != NULL the values and
 * the rhs are produced as well (one NR iteration's stamping). */
typedef struct {
    ivec rows, cols;
    double* vals; int32_t nv;      /* optional values */
    double* rhs;
    int f32;                       /* FP32 path: device values from the FP32 evaluation */
    int ac;                        /* AC admittance: 1 its real part, 2 its imaginary part */
    double omega;                  /* AC angular frequency */
    int32_t n_lin;                 /* stamps of the linear elements (R, V, C), emitted first */
} emitter;
static inline void emit(emitter* e, int32_t r, int32_t c, double val) {
    if (e->vals) { e->vals[e->nv++] = val; }
    else { iv_push(&e->rows, r); iv_push(&e->cols, c); }
}

/* Transient attempt context (restated stepTransient, solvers/circuitsim_solver.cpp:
 * 559-780): step h, companion method (ORC_INT_* = IntegrationType), source
 * values at t + h, capacitor history (branch voltages of the last two
 * accepted steps, current of the last).  NULL for the DC solve. */
typedef struct {
    double h;
    int32_t method;
    const double* vsrc;     /* [n_vsrc] */
    const double* isrc;     /* [n_isrc] */
    const double* hv0;      /* [n_cap]  */
    const double* hv1;      /* [n_cap]  */
    const double* hi0;      /* [n_cap]  */
    int coords_only;        /* pattern build: emit the capacitor coordinates */
} tran_ctx;

static void stamp_all(const nr_netlist* nl, const orc_soa* s, const double* v,
                      double gmin, emitter* e, const tran_ctx* tc) {
    const int32_t n = nl->num_nodes;
    double* rhs = e->rhs;
    /* (1) resistors :2416-2438 */
    for (int32_t i = 0; i < nl->n_res; ++i) {
        int32_t n1 = nl->res_n1[i], n2 = nl->res_n2[i];
        double g = e->ac == 2 ? 0.0 : 1.0 / nl->res_ohms[i];
        if (n1 != 0 && n2 != 0) {
            emit(e, n1 - 1, n1 - 1, g); emit(e, n2 - 1, n2 - 1, g);
            emit(e, n1 - 1, n2 - 1, -g); emit(e, n2 - 1, n1 - 1, -g);
        } else if (n1 != 0) emit(e, n1 - 1, n1 - 1, g);
        else if (n2 != 0) emit(e, n2 - 1, n2 - 1, g);
    }
    /* (2) voltage sources, Norton 1e3 S :2440-2467 */
    for (int32_t i = 0; i < nl->n_vsrc; ++i) {
        int32_t p = nl->vsrc_pos[i], q = nl->vsrc_neg[i];
        double V = (tc && tc->vsrc) ? tc->vsrc[i] : nl->vsrc_volts[i];
        double g = e->ac == 2 ? 0.0 : 1.0e3, ieq = V * g;
        if (p != 0 && q != 0) {
            emit(e, p - 1, p - 1, g); emit(e, q - 1, q - 1, g);
            emit(e, p - 1, q - 1, -g); emit(e, q - 1, p - 1, -g);
        } else if (p != 0) emit(e, p - 1, p - 1, g);
        else if (q != 0) emit(e, q - 1, q - 1, g);
        if (rhs) {
            if (p != 0) rhs[p - 1] += ieq;
            if (q != 0) rhs[q - 1] -= ieq;
        }
    }
    /* (3) current sources :2469-2477 (waveform value in a transient) */
    if (rhs) {
        for (int32_t i = 0; i < nl->n_isrc; ++i) {
            int32_t p = nl->isrc_pos[i], q = nl->isrc_neg[i];
            double I = (tc && tc->isrc) ? tc->isrc[i] : nl->isrc_amps[i];
            if (p != 0) rhs[p - 1] -= I;
            if (q != 0) rhs[q - 1] += I;
        }
    }
    /* (4) capacitors: the companion of the attempt's method (restating
     * :2479-2505 with math/integration_method.h:47-157).  Trapezoidal: G =
     * 2C/h, I_eq = -G v0 - i0, rhs[a] -= I_eq, rhs[b] += I_eq; Gear-1: G =
     * C/h, I_eq = -G v0; Gear-2: G = 1.5C/h, I_src = C (2/h v0 - 0.5/h v1),
     * rhs[a] += I_src, rhs[b] -= I_src.  The history always holds its three
     * slots (netlist/circuit.h:1155-1156), so Gear-2's sparse-history
     * fallback (:131-136) is never taken. */
    for (int32_t i = 0; tc && i < nl->n_cap; ++i) {
        int32_t a = nl->cap_n1[i], b = nl->cap_n2[i];
        const double C = nl->cap_farads[i];
        double G = 0.0, ra = 0.0;   /* ra: added to rhs[a], subtracted from rhs[b] */
        if (e->ac) {
            G = e->ac == 2 ? e->omega * C : 0.0;         /* j omega C (ac_solver.h:160-167) */
        } else if (!tc->coords_only) {
            if (tc->method == ORC_INT_TRAP) {
                G = (2.0 * C) / tc->h;
                ra = -(-G * tc->hv0[i] - tc->hi0[i]);
            } else if (tc->method == ORC_INT_GEAR2) {
                G = (1.5 * C) / tc->h;
                ra = C * ((2.0 / tc->h) * tc->hv0[i] - (0.5 / tc->h) * tc->hv1[i]);
            } else {
                G = C / tc->h;
                ra = -(-G * tc->hv0[i]);
            }
        }
        if (a != 0 && b != 0) {
            emit(e, a - 1, a - 1, G); emit(e, b - 1, b - 1, G);
            emit(e, a - 1, b - 1, -G); emit(e, b - 1, a - 1, -G);
        } else if (a != 0) emit(e, a - 1, a - 1, G);
        else if (b != 0) emit(e, b - 1, b - 1, G);
        if (rhs) {
            if (a != 0) rhs[a - 1] += ra;
            if (b != 0) rhs[b - 1] -= ra;
        }
    }
    e->n_lin = e->vals ? e->nv : e->rows.n;
    /* (6) device evaluation :2535-2537, (7) diodes :2543-2576 */
    for (int32_t i = 0; i < s->nD; ++i) {
        int32_t a = s->da[i], c = s->dc[i];
        double g = 0.0, ieq = 0.0;
        if (e->vals && e->ac) {
            g = e->ac == 2 ? 0.0 : ac_diode_g(volt(v, n, a) - volt(v, n, c), s->dIs[i], s->dN[i],
                                               nl->dio_vt[s->d_src[i]]);
        } else if (e->vals && e->f32) {
            g = diode_g_f32(volt(v, n, a) - volt(v, n, c), s->dIs[i], s->dN[i], s->dVt[i]);
        } else if (e->vals) {
            double vd, id, gd;
            orc_eval_diodes(1, &a, &c, &s->dIs[i], &s->dN[i], &s->dVt[i], v, n, &vd, &id, &gd);
            g = gd; ieq = id - g * vd;
        }
        if (a != 0 && c != 0) {
            emit(e, a - 1, a - 1, g); emit(e, c - 1, c - 1, g);
            emit(e, a - 1, c - 1, -g); emit(e, c - 1, a - 1, -g);
        } else if (a != 0) emit(e, a - 1, a - 1, g);
        else if (c != 0) emit(e, c - 1, c - 1, g);
        if (rhs) {
            if (a != 0) rhs[a - 1] -= ieq;
            if (c != 0) rhs[c - 1] += ieq;
        }
    }
    /* (8) MOSFETs :2592-2657 (gmb = 0 from the generic kernel) */
    for (int32_t i = 0; i < s->nM; ++i) {
        int32_t d = s->md[i], g = s->mg[i], so = s->ms[i], b = s->mb[i];
        double gm = 0, gmb = 0, gds = 0, ieq = 0;
        if (e->vals && e->ac == 2) {
            /* no imaginary part */
        } else if (e->vals && e->f32) {
            const double vs_ = volt(v, n, so);
            float gm32, gds32;
            mos_g_f32(volt(v, n, g) - vs_, volt(v, n, d) - vs_, s->mW[i], s->mL[i], s->mKp[i], s->mVth[i],
                      s->mLam[i], s->mP[i], &gm32, &gds32);
            gm = gm32; gds = gds32;
        } else if (e->vals) {
            double vgs, vds, ids;
            mos_one(volt(v, n, g), volt(v, n, d), volt(v, n, so), s->mW[i], s->mL[i],
                    s->mKp[i], s->mVth[i], s->mLam[i], s->mP[i], &vgs, &vds, &ids, &gm, &gds);
            double vbs = (b != 0) ? (v[b - 1] - (so != 0 ? v[so - 1] : 0.0)) : 0.0;
            ieq = ids - (gm * vgs + gmb * vbs + gds * vds);
        }
        if (d != 0 && so != 0) {
            emit(e, d - 1, d - 1, gds); emit(e, so - 1, so - 1, gds);
            emit(e, d - 1, so - 1, -gds); emit(e, so - 1, d - 1, -gds);
        } else if (d != 0) emit(e, d - 1, d - 1, gds);
        else if (so != 0) emit(e, so - 1, so - 1, gds);
        if (d != 0) {
            if (g != 0) emit(e, d - 1, g - 1, gm);
            if (so != 0) emit(e, d - 1, so - 1, -gm);
        }
        if (so != 0) {
            if (g != 0) emit(e, so - 1, g - 1, -gm);
            emit(e, so - 1, so - 1, gm);
        }
        if (d != 0) {
            if (b != 0) emit(e, d - 1, b - 1, gmb);
            if (so != 0) emit(e, d - 1, so - 1, -gmb);
        }
        if (so != 0) {
            if (b != 0) emit(e, so - 1, b - 1, -gmb);
            emit(e, so - 1, so - 1, gmb);
        }
        if (rhs) {
            if (d != 0) rhs[d - 1] -= ieq;
            if (so != 0) rhs[so - 1] += ieq;
        }
    }
    /* gmin diagonal :2857-2860 */
    for (int32_t i = 0; i < n; ++i) emit(e, i, i, e->ac == 2 ? 0.0 : e->f32 ? (double)(float)gmin : gmin);
}

/* Sparse adjacency with O(1) membership (bitset) for graph elimination. */
typedef struct {
    int32_t nv;            /* vertices 0..nv-1 */
    int64_t words;         /* per-row words */
    uint64_t* bits;
    ivec* nb;
} egraph;
static void eg_init(egraph* g, int32_t nv) {
    g->nv = nv;
    g->words = (nv + 63) / 64;
    g->bits = (uint64_t*)xcalloc((size_t)(g->words * nv), 8);
    g->nb = (ivec*)xcalloc((size_t)nv, sizeof(ivec));
}
static void eg_free(egraph* g) {
    for (int32_t i = 0; i < g->nv; ++i) free(g->nb[i].a);
    free(g->nb); free(g->bits);
}
static inline int eg_has(const egraph* g, int32_t u, int32_t w) {
    return (int)((g->bits[(int64_t)u * g->words + (w >> 6)] >> (w & 63)) & 1u);
}
static inline void eg_add(egraph* g, int32_t u, int32_t w) {
    if (u == w || eg_has(g, u, w)) return;
    g->bits[(int64_t)u * g->words + (w >> 6)] |= (uint64_t)1 << (w & 63);
    g->bits[(int64_t)w * g->words + (u >> 6)] |= (uint64_t)1 << (u & 63);
    iv_push(&g->nb[u], w); iv_push(&g->nb[w], u);
}
static void eg_remove_vertex(egraph* g, int32_t v) {
    ivec* nv = &g->nb[v];
    for (int32_t k = 0; k < nv->n; ++k) {
        int32_t u = nv->a[k];
        ivec* nu = &g->nb[u];
        for (int32_t j = 0; j < nu->n; ++j)
            if (nu->a[j] == v) { nu->a[j] = nu->a[--nu->n]; break; }
        g->bits[(int64_t)u * g->words + (v >> 6)] &= ~((uint64_t)1 << (v & 63));
        g->bits[(int64_t)v * g->words + (u >> 6)] &= ~((uint64_t)1 << (u & 63));
    }
    nv->n = 0;
}
static void eg_cliquify(egraph* g, int32_t v) {
    ivec* nv = &g->nb[v];
    int32_t cnt = nv->n;
    int32_t* tmp = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(cnt + 1));
    memcpy(tmp, nv->a, sizeof(int32_t) * (size_t)cnt);
    for (int32_t i = 0; i < cnt; ++i)
        for (int32_t j = i + 1; j < cnt; ++j) eg_add(g, tmp[i], tmp[j]);
    free(tmp);
}
static uint32_t eg_fill(const egraph* g, int32_t v) {
    const ivec* nv = &g->nb[v];
    uint32_t f = 0;
    for (int32_t i = 0; i < nv->n; ++i)
        for (int32_t j = i + 1; j < nv->n; ++j)
            if (!eg_has(g, nv->a[i], nv->a[j])) ++f;
    return f;
}
static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* a9 + a10: FactorGraph::fromTensorizedBlock (infrastructure/factor_graph.cpp
 * :9-162) and TreewidthAnalyzer::analyze (infrastructure/treewidth_analyzer.cpp
 * :14-141), restated literally: every step scans all remaining variables in
 * ascending id order, fill = number of non-adjacent neighbour pairs, argmin
 * with ties to the smallest id; branch-current ids start at numCircuitNodes
 * and may collide with a node id (kept as coded). */
static void min_fill(const nr_netlist* nl, const orc_soa* s, orc_topo* t) {
    const int32_t ncn = nl->num_nodes;
    const int32_t nbranch = nl->n_vsrc;        /* + inductors (none in this ABI) */
    const int32_t maxid = ncn + nbranch;       /* ids 0..maxid */
    egraph g;
    eg_init(&g, maxid + 1);
    char* touched = (char*)xcalloc((size_t)maxid + 1, 1);
    int32_t branch = ncn;
#define FACTOR_BEGIN() int32_t fv[8]; int32_t nfv = 0
#define FACTOR_ADD(x) do { if ((x) > 0) { fv[nfv++] = (x); touched[(x)] = 1; } } while (0)
#define FACTOR_END() do { for (int32_t i_ = 0; i_ < nfv; ++i_) for (int32_t j_ = i_ + 1; j_ < nfv; ++j_) \
        if (fv[i_] != fv[j_]) eg_add(&g, fv[i_], fv[j_]); } while (0)
    for (int32_t i = 0; i < s->nD; ++i) { FACTOR_BEGIN(); FACTOR_ADD(s->da[i]); FACTOR_ADD(s->dc[i]); FACTOR_END(); }
    for (int32_t i = 0; i < s->nM; ++i) {
        FACTOR_BEGIN(); FACTOR_ADD(s->md[i]); FACTOR_ADD(s->mg[i]); FACTOR_ADD(s->ms[i]); FACTOR_ADD(s->mb[i]); FACTOR_END();
    }
    for (int32_t i = 0; i < nl->n_res; ++i) { FACTOR_BEGIN(); FACTOR_ADD(nl->res_n1[i]); FACTOR_ADD(nl->res_n2[i]); FACTOR_END(); }
    for (int32_t i = 0; i < nl->n_cap; ++i) { FACTOR_BEGIN(); FACTOR_ADD(nl->cap_n1[i]); FACTOR_ADD(nl->cap_n2[i]); FACTOR_END(); }
    for (int32_t i = 0; i < nl->n_vsrc; ++i) {
        FACTOR_BEGIN(); FACTOR_ADD(nl->vsrc_pos[i]); FACTOR_ADD(nl->vsrc_neg[i]);
        fv[nfv++] = branch++;            /* branch-current variable */
        FACTOR_END();
    }
    for (int32_t i = 0; i < nl->n_isrc; ++i) { FACTOR_BEGIN(); FACTOR_ADD(nl->isrc_pos[i]); FACTOR_ADD(nl->isrc_neg[i]); FACTOR_END(); }
#undef FACTOR_BEGIN
#undef FACTOR_ADD
#undef FACTOR_END
    /* variables: sorted touched node ids, then branch ids (possibly colliding) */
    int32_t nvars = 0;
    int32_t* rem = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(maxid + nbranch + 2));
    for (int32_t id = 1; id <= maxid; ++id)
        if (id <= ncn && touched[id]) rem[nvars++] = id;
    for (int32_t k = 0; k < nbranch; ++k) rem[nvars++] = ncn + k;
    free(touched);
    qsort(rem, (size_t)nvars, sizeof(int32_t), cmp_i32);
    t->n_vars = nvars;
    t->ordering_len = nvars;
    t->ordering_raw = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(nvars + 1));
    int32_t nrem = nvars;
    uint32_t maxClique = 0;
    for (int32_t step = 0; step < nvars; ++step) {
        int32_t best = rem[0], bestpos = 0;
        uint32_t bestFill = UINT32_MAX;
        for (int32_t k = 0; k < nrem; ++k) {
            int32_t v = rem[k];
            uint32_t f = eg_fill(&g, v);
            if (f < bestFill || (f == bestFill && v < best)) { bestFill = f; best = v; bestpos = k; }
        }
        /* std::find returns the first copy of `best` (rem is sorted) */
        for (int32_t k = 0; k < nrem; ++k) if (rem[k] == best) { bestpos = k; break; }
        t->ordering_raw[step] = best;
        uint32_t clique = (uint32_t)g.nb[best].n + 1;
        if (clique > maxClique) maxClique = clique;
        eg_cliquify(&g, best);
        eg_remove_vertex(&g, best);
        memmove(rem + bestpos, rem + bestpos + 1, sizeof(int32_t) * (size_t)(nrem - bestpos - 1));
        --nrem;
    }
    t->treewidth = maxClique > 0 ? (int32_t)maxClique - 1 : 0;
    free(rem);
    eg_free(&g);
}

/* Projection of the Min-Fill ordering onto matrix rows: keep the first
 * occurrence of each node id 1..n, drop ids above n, row = id - 1; rows that
 * no device touches (none in practice) are appended in ascending order. */
static void project_ordering(orc_topo* t) {
    const int32_t n = t->n;
    char* seen = (char*)xcalloc((size_t)n + 1, 1);
    int32_t k = 0;
    t->perm = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(n + 1));
    t->iperm = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(n + 1));
    for (int32_t i = 0; i < t->ordering_len; ++i) {
        int32_t id = t->ordering_raw[i];
        if (id >= 1 && id <= n && !seen[id]) { seen[id] = 1; t->perm[k++] = id - 1; }
    }
    for (int32_t id = 1; id <= n; ++id) if (!seen[id]) t->perm[k++] = id - 1;
    for (int32_t i = 0; i < n; ++i) t->iperm[t->perm[i]] = i;
    free(seen);
}

/* a11: symbolicFactorize (math/sparse_lu.cpp:14-128) restated. */
static void symbolic(orc_topo* t) {
    const int32_t n = t->n;
    egraph g;
    eg_init(&g, n);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t k = t->row_ptr[r]; k < t->row_ptr[r + 1]; ++k)
            if (t->col_idx[k] != r) eg_add(&g, r, t->col_idx[k]);
    ivec* rows = (ivec*)xcalloc((size_t)n, sizeof(ivec));
    for (int32_t i = 0; i < n; ++i) iv_push(&rows[i], i);
    int32_t fill = 0;
    for (int32_t step = 0; step < n; ++step) {
        int32_t v = t->perm[step];
        ivec* nv = &g.nb[v];
        for (int32_t k = 0; k < nv->n; ++k) {
            int32_t pc = t->iperm[nv->a[k]];
            iv_push(&rows[step], pc);
            iv_push(&rows[pc], step);
        }
        int32_t cnt = nv->n;
        int32_t* tmp = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(cnt + 1));
        memcpy(tmp, nv->a, sizeof(int32_t) * (size_t)cnt);
        for (int32_t i = 0; i < cnt; ++i)
            for (int32_t j = i + 1; j < cnt; ++j)
                if (!eg_has(&g, tmp[i], tmp[j])) { eg_add(&g, tmp[i], tmp[j]); ++fill; }
        free(tmp);
        eg_remove_vertex(&g, v);
    }
    eg_free(&g);
    t->fill_in = fill;
    t->frow_ptr = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(n + 1));
    int64_t tot = 0;
    for (int32_t i = 0; i < n; ++i) {
        qsort(rows[i].a, (size_t)rows[i].n, sizeof(int32_t), cmp_i32);
        int32_t u = 0;   /* dedupe (the reference uses a set per row) */
        for (int32_t k = 0; k < rows[i].n; ++k)
            if (k == 0 || rows[i].a[k] != rows[i].a[k - 1]) rows[i].a[u++] = rows[i].a[k];
        rows[i].n = u;
        tot += u;
    }
    t->nnz_f = (int32_t)tot;
    t->fcol_idx = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(tot + 1));
    t->fdiag = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t p = 0;
    for (int32_t i = 0; i < n; ++i) {
        t->frow_ptr[i] = p;
        for (int32_t k = 0; k < rows[i].n; ++k) {
            if (rows[i].a[k] == i) t->fdiag[i] = p;
            t->fcol_idx[p++] = rows[i].a[k];
        }
        free(rows[i].a);
    }
    t->frow_ptr[n] = p;
    free(rows);
}

/* Relaxed supernodes of the filled pattern (see supernodes()): consecutive
 * columns along an elimination-tree chain.  Children are listed in
 * increasing supernode index; the front of supernode J is ordered [own
 * columns, then its update rows R = struct L(:, last column)]. */
static int32_t front_local(const orc_topo* t, int32_t J, int32_t x) {
    int32_t f = t->sn_first[J], c = t->sn_ncols[J];
    if (x >= f && x < f + c) return x - f;
    const int32_t* R = t->sn_rows + t->sn_row_off[J];
    int32_t lo = 0, hi = t->sn_nrows[J];
    while (lo < hi) { int32_t mid = (lo + hi) / 2; if (R[mid] < x) lo = mid + 1; else hi = mid; }
    return (lo < t->sn_nrows[J] && R[lo] == x) ? c + lo : -1;
}

static void supernodes(orc_topo* t) {
    const int32_t n = t->n;
    int32_t* cnt = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);     /* |struct L(:,j)| off-diagonal */
    int32_t* par = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)n);
    int32_t* nch = (int32_t*)xcalloc((size_t)n, 4);
    for (int32_t j = 0; j < n; ++j) {
        cnt[j] = t->frow_ptr[j + 1] - t->fdiag[j] - 1;
        par[j] = cnt[j] > 0 ? t->fcol_idx[t->fdiag[j] + 1] : -1;
        if (par[j] >= 0) nch[par[j]]++;
    }
    /* relaxed supernodes: from the last column down, column j-1 joins the
     * supernode [j, last] when par(j-1) = j and either the pair is
     * fundamental (no explicit zero, j has no other child) or the merged
     * front's explicit zeros stay within a fifth of its L entries at front
     * order <= 512 and at most 16 pivots */
    int32_t* start = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(n + 1));
    char* begins = (char*)xcalloc((size_t)n + 1, 1);
    {
        int32_t last = n - 1;
        int64_t zeros = 0;
        for (int32_t j = n - 1; j >= 1; --j) {
            int join = 0;
            if (par[j - 1] == j) {
                const int64_t cl = cnt[last];
                const int64_t c2 = last - j + 2, r2 = c2 + cl;
                const int64_t zadd = (last - j + 1 + cl) - cnt[j - 1];
                const int64_t ents = c2 * r2 - c2 * (c2 + 1) / 2;
                if (zadd == 0 && nch[j] == 1) join = 1;
                else if (r2 <= 512 && c2 <= 16 && 5 * (zeros + zadd) <= ents) { join = 1; zeros += zadd; }
            }
            if (!join) { begins[j] = 1; last = j - 1; zeros = 0; }
        }
    }
    int32_t ns = 0;
    start[ns++] = 0;
    for (int32_t j = 1; j < n; ++j)
        if (begins[j]) start[ns++] = j;
    start[ns] = n;
    free(begins);
    t->ns = ns;
    t->sn_first = (int32_t*)xmalloc(4 * (size_t)ns); t->sn_ncols = (int32_t*)xmalloc(4 * (size_t)ns);
    t->sn_nrows = (int32_t*)xmalloc(4 * (size_t)ns); t->sn_row_off = (int32_t*)xmalloc(4 * (size_t)ns);
    int32_t* col2sn = (int32_t*)xmalloc(4 * (size_t)n);
    int64_t rtot = 0;
    for (int32_t J = 0; J < ns; ++J) {
        t->sn_first[J] = start[J];
        t->sn_ncols[J] = start[J + 1] - start[J];
        t->sn_nrows[J] = cnt[start[J + 1] - 1];
        t->sn_row_off[J] = (int32_t)rtot;
        rtot += t->sn_nrows[J];
        for (int32_t j = start[J]; j < start[J + 1]; ++j) col2sn[j] = J;
    }
    t->sn_rows = (int32_t*)xmalloc(4 * (size_t)(rtot + 1));
    for (int32_t J = 0; J < ns; ++J) {
        int32_t last = start[J + 1] - 1;
        memcpy(t->sn_rows + t->sn_row_off[J], t->fcol_idx + t->fdiag[last] + 1, 4 * (size_t)t->sn_nrows[J]);
    }
    /* children in increasing index */
    t->sn_child_ptr = (int32_t*)xcalloc((size_t)ns + 1, 4);
    int32_t* snpar = (int32_t*)xmalloc(4 * (size_t)ns);
    for (int32_t J = 0; J < ns; ++J) {
        int32_t last = start[J + 1] - 1;
        snpar[J] = par[last] >= 0 ? col2sn[par[last]] : -1;
        if (snpar[J] >= 0) t->sn_child_ptr[snpar[J] + 1]++;
    }
    for (int32_t J = 0; J < ns; ++J) t->sn_child_ptr[J + 1] += t->sn_child_ptr[J];
    t->sn_child = (int32_t*)xmalloc(4 * (size_t)(ns + 1));
    int32_t* fillp = (int32_t*)xmalloc(4 * (size_t)(ns + 1));
    memcpy(fillp, t->sn_child_ptr, 4 * (size_t)ns);
    for (int32_t J = 0; J < ns; ++J) if (snpar[J] >= 0) t->sn_child[fillp[snpar[J]]++] = J;
    /* A entries per supernode (CSR order) */
    t->sn_a_ptr = (int32_t*)xcalloc((size_t)ns + 1, 4);
    t->sn_a_slot = (int32_t*)xmalloc(4 * (size_t)(t->nnz_a + 1));
    t->sn_a_li = (int32_t*)xmalloc(4 * (size_t)(t->nnz_a + 1));
    t->sn_a_lj = (int32_t*)xmalloc(4 * (size_t)(t->nnz_a + 1));
    int32_t* owner = (int32_t*)xmalloc(4 * (size_t)(t->nnz_a + 1));
    for (int32_t r = 0; r < n; ++r)
        for (int32_t k = t->row_ptr[r]; k < t->row_ptr[r + 1]; ++k) {
            int32_t pi = t->iperm[r], pj = t->iperm[t->col_idx[k]];
            owner[k] = col2sn[pi < pj ? pi : pj];
            t->sn_a_ptr[owner[k] + 1]++;
        }
    for (int32_t J = 0; J < ns; ++J) t->sn_a_ptr[J + 1] += t->sn_a_ptr[J];
    memcpy(fillp, t->sn_a_ptr, 4 * (size_t)ns);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t k = t->row_ptr[r]; k < t->row_ptr[r + 1]; ++k) {
            int32_t J = owner[k], q = fillp[J]++;
            t->sn_a_slot[q] = k;
            t->sn_a_li[q] = front_local(t, J, t->iperm[r]);
            t->sn_a_lj[q] = front_local(t, J, t->iperm[t->col_idx[k]]);
        }
    free(owner); free(fillp); free(snpar); free(col2sn); free(start); free(cnt); free(par); free(nch);
}

/* a8: createCsr / buildPattern (math/linalg.h:281-431): sort the triplet
 * coordinates by (row, col), merge duplicates; the stamp map is each
 * triplet's merged slot (assembleFast's position_map lookup, :441-465). */
static _Thread_local int64_t* g_sort_keys;   /* per thread: topologies may be built concurrently */
static int cmp_key_idx(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    int64_t kx = g_sort_keys[x], ky = g_sort_keys[y];
    if (kx != ky) return (kx > ky) - (kx < ky);
    return (x > y) - (x < y);
}

int orc_topology_build(const nr_netlist* nl, orc_topo** out) { return orc_topology_build_ex(nl, 0, out); }

static int topology_build(const nr_netlist* nl, int32_t transient, const int32_t* ordering, int32_t ordering_len,
                          orc_topo** out);

int orc_topology_build_ex(const nr_netlist* nl, int32_t transient, orc_topo** out) {
    return topology_build(nl, transient, NULL, 0, out);
}

/* The same topology with the raw Min-Fill ordering given (an earlier run's
 * NR_EXPORT_ORDERING_RAW of the same netlist), so that the literal Min-Fill,
 * O(n^2 d^2), need not run again.  The treewidth is taken from the fill:
 * the largest number of entries right of the diagonal in a row of U. */
int orc_topology_build_ordered(const nr_netlist* nl, int32_t transient, const int32_t* ordering,
                               int32_t ordering_len, orc_topo** out) {
    if (!ordering || ordering_len <= 0) return NR_ERR_INVALID;
    return topology_build(nl, transient, ordering, ordering_len, out);
}

static int topology_build(const nr_netlist* nl, int32_t transient, const int32_t* ordering, int32_t ordering_len,
                          orc_topo** out) {
    if (!nl || nl->num_nodes <= 0) return NR_ERR_INVALID;
    orc_topo* t = (orc_topo*)xcalloc(1, sizeof(orc_topo));
    t->n = nl->num_nodes;
    t->transient = transient ? 1 : 0;
    t->soa = orc_tensorize(nl);
    emitter e;
    memset(&e, 0, sizeof(e));
    tran_ctx coords = {0.0, ORC_INT_GEAR1, NULL, NULL, NULL, NULL, NULL, 1};
    stamp_all(nl, t->soa, NULL, 0.0, &e, transient ? &coords : NULL);
    t->n_stamps = e.rows.n;
    t->st_row = e.rows.a; t->st_col = e.cols.a;
    const int32_t S = t->n_stamps, n = t->n;
    int64_t* keys = (int64_t*)xmalloc(sizeof(int64_t) * (size_t)S);
    int32_t* idx = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)S);
    for (int32_t k = 0; k < S; ++k) { keys[k] = ((int64_t)t->st_row[k] << 32) | (uint32_t)t->st_col[k]; idx[k] = k; }
    g_sort_keys = keys;
    qsort(idx, (size_t)S, sizeof(int32_t), cmp_key_idx);
    g_sort_keys = NULL;
    t->stamp_slot = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)S);
    t->row_ptr = (int32_t*)xcalloc((size_t)n + 1, 4);
    t->col_idx = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)S);
    int32_t nnz = 0;
    int64_t last = -1;
    for (int32_t k = 0; k < S; ++k) {
        int32_t j = idx[k];
        if (keys[j] != last) {
            t->col_idx[nnz] = t->st_col[j];
            t->row_ptr[t->st_row[j] + 1]++;
            ++nnz;
            last = keys[j];
        }
        t->stamp_slot[j] = nnz - 1;
    }
    for (int32_t i = 0; i < n; ++i) t->row_ptr[i + 1] += t->row_ptr[i];
    t->nnz_a = nnz;
    free(keys); free(idx);
    if (ordering) {
        t->ordering_len = t->n_vars = ordering_len;
        t->ordering_raw = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(ordering_len + 1));
        memcpy(t->ordering_raw, ordering, sizeof(int32_t) * (size_t)ordering_len);
    } else {
        min_fill(nl, t->soa, t);
    }
    project_ordering(t);
    symbolic(t);
    if (ordering) {
        int32_t tw = 0;
        for (int32_t i = 0; i < n; ++i) {
            const int32_t right = t->frow_ptr[i + 1] - t->fdiag[i] - 1;
            if (right > tw) tw = right;
        }
        t->treewidth = tw;
    }
    supernodes(t);
    *out = t;
    return NR_OK;
}

void orc_topology_free(orc_topo* t) {
    if (!t) return;
    orc_soa_free(t->soa);
    free(t->st_row); free(t->st_col); free(t->row_ptr); free(t->col_idx); free(t->stamp_slot);
    free(t->ordering_raw); free(t->perm); free(t->iperm);
    free(t->frow_ptr); free(t->fcol_idx); free(t->fdiag);
    free(t->sn_first); free(t->sn_ncols); free(t->sn_nrows); free(t->sn_row_off); free(t->sn_rows);
    free(t->sn_child_ptr); free(t->sn_child);
    free(t->sn_a_ptr); free(t->sn_a_slot); free(t->sn_a_li); free(t->sn_a_lj);
    free(t);
}

void orc_topology_info(const orc_topo* t, int32_t* info) {
    info[0] = t->n; info[1] = t->nnz_a; info[2] = t->nnz_f; info[3] = t->n_stamps;
    info[4] = t->treewidth; info[5] = t->ordering_len; info[6] = t->fill_in; info[7] = t->n_vars;
}

int64_t orc_topology_export(const orc_topo* t, int32_t what, int32_t* out, int64_t cap) {
    const int32_t* src = NULL;
    int64_t cnt = 0;
    switch (what) {
    case NR_EXPORT_ROW_PTR: src = t->row_ptr; cnt = t->n + 1; break;
    case NR_EXPORT_COL_IDX: src = t->col_idx; cnt = t->nnz_a; break;
    case NR_EXPORT_STAMP_SLOT: src = t->stamp_slot; cnt = t->n_stamps; break;
    case NR_EXPORT_ORDERING_RAW: src = t->ordering_raw; cnt = t->ordering_len; break;
    case NR_EXPORT_PERM: src = t->perm; cnt = t->n; break;
    case NR_EXPORT_FILL_ROW_PTR: src = t->frow_ptr; cnt = t->n + 1; break;
    case NR_EXPORT_FILL_COL_IDX: src = t->fcol_idx; cnt = t->nnz_f; break;
    case NR_EXPORT_DIODE_ORDER: src = t->soa->d_src; cnt = t->soa->nD; break;
    case NR_EXPORT_MOS_ORDER: src = t->soa->m_src; cnt = t->soa->nM; break;
    case NR_EXPORT_SN_FIRST: src = t->sn_first; cnt = t->ns; break;
    case ORC_EXPORT_STAMP_ROW: src = t->st_row; cnt = t->n_stamps; break;
    case ORC_EXPORT_STAMP_COL: src = t->st_col; cnt = t->n_stamps; break;
    default: return -1;
    }
    if (cnt > cap) return -1;
    if (cnt) memcpy(out, src, sizeof(int32_t) * (size_t)cnt);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* stamping + assembly of one iteration                                     */
/* ------------------------------------------------------------------------ */
static int stamp_with_soa(const nr_netlist* nl, const orc_soa* s, const orc_topo* t,
                          const double* v, double gmin, double* vals, double* rhs, const tran_ctx* tc) {
    emitter e;
    memset(&e, 0, sizeof(e));
    e.vals = vals;
    e.rhs = rhs;
    memset(rhs, 0, sizeof(double) * (size_t)nl->num_nodes);
    stamp_all(nl, s, v, gmin, &e, tc);
    return e.nv == t->n_stamps ? NR_OK : NR_ERR_TOPOLOGY;
}

int orc_stamp(const nr_netlist* nl, const orc_topo* t, const double* v,
              double gmin, double* vals, double* rhs) {
    orc_soa* s = orc_tensorize(nl);
    int rc = stamp_with_soa(nl, s, t, v, gmin, vals, rhs, NULL);
    orc_soa_free(s);
    return rc;
}

int orc_assemble(const orc_topo* t, const double* vals, double* a_vals) {
    memset(a_vals, 0, sizeof(double) * (size_t)t->nnz_a);
    for (int32_t k = 0; k < t->n_stamps; ++k) a_vals[t->stamp_slot[k]] += vals[k];
    return NR_OK;
}

/* FP32 path: the Jacobian in float.  The linear elements' stamps (emitted
 * first) are summed per slot in FP64 and rounded once; the FP32 device
 * stamps and the gmin diagonal are then added in float, in emission order. */
static int assemble_f32(const nr_netlist* nl, const orc_soa* s, const orc_topo* t, const double* v, double gmin,
                        const tran_ctx* tc, double* vals, double* lin, float* a32) {
    emitter e;
    memset(&e, 0, sizeof(e));
    e.vals = vals;
    e.f32 = 1;
    stamp_all(nl, s, v, gmin, &e, tc);
    if (e.nv != t->n_stamps) return NR_ERR_TOPOLOGY;
    for (int32_t q = 0; q < t->nnz_a; ++q) lin[q] = 0.0;
    for (int32_t k = 0; k < e.n_lin; ++k) lin[t->stamp_slot[k]] += vals[k];
    for (int32_t q = 0; q < t->nnz_a; ++q) a32[q] = (float)lin[q];
    for (int32_t k = e.n_lin; k < t->n_stamps; ++k) a32[t->stamp_slot[k]] = a32[t->stamp_slot[k]] + (float)vals[k];
    return NR_OK;
}

/* ------------------------------------------------------------------------ */
/* a12: linear solves                                                       */
/* ------------------------------------------------------------------------ */
int orc_lu_dense(int32_t n, const int32_t* rp, const int32_t* ci,
                 const double* vals, const double* rhs, double* x,
                 double* smallest_pivot) {
    double* A = (double*)xcalloc((size_t)n * (size_t)n, 8);
    double* b = (double*)xmalloc(sizeof(double) * (size_t)n);
    for (int32_t r = 0; r < n; ++r)
        for (int32_t k = rp[r]; k < rp[r + 1]; ++k) A[(size_t)r * n + ci[k]] += vals[k];
    memcpy(b, rhs, sizeof(double) * (size_t)n);
    double smallest = 1e300;
    for (int32_t i = 0; i < n; ++i) {
        double maxVal = 0.0;
        int32_t maxRow = i;
        for (int32_t k = i; k < n; ++k) {
            double val = fabs(A[(size_t)k * n + i]);
            if (val > maxVal) { maxVal = val; maxRow = k; }
            else if (val == maxVal && k < maxRow) maxRow = k;
        }
        if (maxVal < smallest) smallest = maxVal;
        if (maxVal < 1e-25) {
            for (int32_t k = 0; k < n; ++k) x[k] = 0.0;
            free(A); free(b);
            if (smallest_pivot) *smallest_pivot = smallest;
            return NR_STATUS_SINGULAR;
        }
        if (maxRow != i) {
            for (int32_t k = 0; k < n; ++k) {
                double tmp = A[(size_t)i * n + k];
                A[(size_t)i * n + k] = A[(size_t)maxRow * n + k];
                A[(size_t)maxRow * n + k] = tmp;
            }
            double tb = b[i]; b[i] = b[maxRow]; b[maxRow] = tb;
        }
        const double piv = A[(size_t)i * n + i];
        for (int32_t j = i + 1; j < n; ++j) {
            double* rj = &A[(size_t)j * n];
            rj[i] /= piv;
            const double l = rj[i];
            const double* ri = &A[(size_t)i * n];
            if (l != 0.0)
                for (int32_t k = i + 1; k < n; ++k) rj[k] -= l * ri[k];
        }
    }
    for (int32_t i = 0; i < n; ++i) {
        double y = b[i];
        for (int32_t j = 0; j < i; ++j) y -= A[(size_t)i * n + j] * b[j];
        b[i] = y;
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double xi = b[i];
        for (int32_t j = i + 1; j < n; ++j) xi -= A[(size_t)i * n + j] * x[j];
        x[i] = xi / A[(size_t)i * n + i];
    }
    if (smallest_pivot) *smallest_pivot = smallest;
    free(A); free(b);
    return NR_OK;
}

/* Row-oriented (IKJ) LU without pivoting on the filled pattern in permuted
 * space; the pattern is closed under elimination, so every update lands on
 * a pattern position. */
int orc_lu_sparse(const orc_topo* t, const double* a_vals, const double* rhs,
                  double* x, double* smallest_pivot) {
    const int32_t n = t->n;
    const int32_t *fr = t->frow_ptr, *fc = t->fcol_idx, *fd = t->fdiag;
    double* F = (double*)xcalloc((size_t)t->nnz_f, 8);
    double* w = (double*)xcalloc((size_t)n, 8);
    double* y = (double*)xmalloc(sizeof(double) * (size_t)n);
    /* scatter A into the permuted pattern */
    for (int32_t r = 0; r < n; ++r) {
        int32_t pr = t->iperm[r];
        for (int32_t k = t->row_ptr[r]; k < t->row_ptr[r + 1]; ++k) {
            int32_t pc = t->iperm[t->col_idx[k]];
            w[pc] = a_vals[k];
        }
        for (int32_t k = fr[pr]; k < fr[pr + 1]; ++k) { F[k] = w[fc[k]]; }
        for (int32_t k = t->row_ptr[r]; k < t->row_ptr[r + 1]; ++k) w[t->iperm[t->col_idx[k]]] = 0.0;
    }
    double smallest = 1e300;
    int rc = NR_OK;
    for (int32_t i = 0; i < n && rc == NR_OK; ++i) {
        for (int32_t k = fr[i]; k < fr[i + 1]; ++k) w[fc[k]] = F[k];
        for (int32_t k = fr[i]; k < fd[i]; ++k) {
            int32_t j = fc[k];
            double l = w[j] / F[fd[j]];
            w[j] = l;
            for (int32_t q = fd[j] + 1; q < fr[j + 1]; ++q) w[fc[q]] -= l * F[q];
        }
        for (int32_t k = fr[i]; k < fr[i + 1]; ++k) { F[k] = w[fc[k]]; w[fc[k]] = 0.0; }
        double piv = fabs(F[fd[i]]);
        if (piv < smallest) smallest = piv;
        if (!(piv >= 1e-25)) rc = NR_STATUS_SINGULAR;
    }
    if (rc == NR_OK) {
        for (int32_t i = 0; i < n; ++i) {
            double s = rhs[t->perm[i]];
            for (int32_t k = fr[i]; k < fd[i]; ++k) s -= F[k] * y[fc[k]];
            y[i] = s;
        }
        for (int32_t i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int32_t k = fd[i] + 1; k < fr[i + 1]; ++k) s -= F[k] * y[fc[k]];
            y[i] = s / F[fd[i]];
        }
        for (int32_t i = 0; i < n; ++i) x[t->perm[i]] = y[i];
    } else {
        for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
    }
    if (smallest_pivot) *smallest_pivot = smallest;
    free(F); free(w); free(y);
    return rc;
}

int orc_lu_multifrontal(const orc_topo* t, const double* a_vals, const double* rhs,
                        double* x, double* smallest_pivot) {
    const int32_t n = t->n, ns = t->ns;
    int64_t* u_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t* cb_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t* rcb_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t ut = 0, ct = 0, rt = 0;
    int32_t rmax = 1;
    for (int32_t J = 0; J < ns; ++J) {
        int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m;
        u_off[J] = ut; ut += (int64_t)c * r;
        cb_off[J] = ct; ct += (int64_t)m * m;
        rcb_off[J] = rt; rt += m;
        if (r > rmax) rmax = r;
    }
    double* U = (double*)xmalloc(8 * (size_t)(ut + 1));
    double* CB = (double*)xmalloc(8 * (size_t)(ct + 1));
    double* RCB = (double*)xmalloc(8 * (size_t)(rt + 1));
    double* y = (double*)xmalloc(8 * (size_t)n);
    double* xp = (double*)xmalloc(8 * (size_t)n);
    double* F = (double*)xmalloc(8 * (size_t)rmax * (size_t)rmax);
    double* fr = (double*)xmalloc(8 * (size_t)rmax);
    double smallest = 1e300;
    int rc = NR_OK;
    for (int32_t J = 0; J < ns; ++J) {
        const int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m, f = t->sn_first[J];
        for (int32_t e = 0; e < r * r; ++e) F[e] = 0.0;
        for (int32_t e = 0; e < r; ++e) fr[e] = 0.0;
        for (int32_t q = t->sn_a_ptr[J]; q < t->sn_a_ptr[J + 1]; ++q)
            F[t->sn_a_li[q] * r + t->sn_a_lj[q]] += a_vals[t->sn_a_slot[q]];
        for (int32_t k = 0; k < c; ++k) fr[k] += rhs[t->perm[f + k]];
        for (int32_t q = t->sn_child_ptr[J]; q < t->sn_child_ptr[J + 1]; ++q) {
            const int32_t C = t->sn_child[q], mc = t->sn_nrows[C];
            const int32_t* Rc = t->sn_rows + t->sn_row_off[C];
            int32_t* rel = (int32_t*)xmalloc(4 * (size_t)(mc + 1));
            for (int32_t a = 0; a < mc; ++a) rel[a] = front_local(t, J, Rc[a]);
            const double* cb = CB + cb_off[C];
            for (int32_t a = 0; a < mc; ++a)
                for (int32_t b = 0; b < mc; ++b) F[rel[a] * r + rel[b]] += cb[a * mc + b];
            for (int32_t a = 0; a < mc; ++a) fr[rel[a]] += RCB[rcb_off[C] + a];
            free(rel);
        }
        for (int32_t k = 0; k < c; ++k) {
            const double piv = F[k * r + k];
            if (fabs(piv) < smallest) smallest = fabs(piv);
            if (!(fabs(piv) >= 1e-25)) rc = NR_STATUS_SINGULAR;
            for (int32_t i = k + 1; i < r; ++i) F[i * r + k] = F[i * r + k] / piv;
            for (int32_t i = k + 1; i < r; ++i) {
                const double l = F[i * r + k];
                for (int32_t j = k + 1; j < r; ++j) F[i * r + j] = fma(-l, F[k * r + j], F[i * r + j]);
                fr[i] = fma(-l, fr[k], fr[i]);
            }
        }
        memcpy(U + u_off[J], F, 8 * (size_t)c * (size_t)r);
        for (int32_t k = 0; k < c; ++k) y[f + k] = fr[k];
        for (int32_t a = 0; a < m; ++a) {
            for (int32_t b = 0; b < m; ++b) CB[cb_off[J] + a * m + b] = F[(c + a) * r + c + b];
            RCB[rcb_off[J] + a] = fr[c + a];
        }
    }
    if (rc == NR_OK) {
        for (int32_t J = ns - 1; J >= 0; --J) {
            const int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m, f = t->sn_first[J];
            const double* Uj = U + u_off[J];
            const int32_t* R = t->sn_rows + t->sn_row_off[J];
            for (int32_t k = 0; k < c; ++k) {
                double s = y[f + k];
                for (int32_t a = 0; a < m; ++a) s = fma(-Uj[k * r + c + a], xp[R[a]], s);
                fr[k] = s;
            }
            for (int32_t k = c - 1; k >= 0; --k) {
                const double xk = fr[k] / Uj[k * r + k];
                xp[f + k] = xk;
                for (int32_t i = 0; i < k; ++i) fr[i] = fma(-Uj[i * r + k], xk, fr[i]);
            }
        }
        for (int32_t i = 0; i < n; ++i) x[t->perm[i]] = xp[i];
    } else {
        for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
    }
    if (smallest_pivot) *smallest_pivot = smallest;
    free(u_off); free(cb_off); free(rcb_off); free(U); free(CB); free(RCB); free(y); free(xp); free(F); free(fr);
    return rc;
}

/* The same multifrontal LU in single precision (float storage, float FMAs,
 * the same operation order), for the FP32 path. */
int orc_lu_multifrontal_f32(const orc_topo* t, const float* a_vals, const float* rhs, float* x) {
    const int32_t n = t->n, ns = t->ns;
    int64_t* u_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t* cb_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t* rcb_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t ut = 0, ct = 0, rt = 0;
    int32_t rmax = 1;
    for (int32_t J = 0; J < ns; ++J) {
        int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m;
        u_off[J] = ut; ut += (int64_t)c * r;
        cb_off[J] = ct; ct += (int64_t)m * m;
        rcb_off[J] = rt; rt += m;
        if (r > rmax) rmax = r;
    }
    float* U = (float*)xmalloc(4 * (size_t)(ut + 1));
    float* CB = (float*)xmalloc(4 * (size_t)(ct + 1));
    float* RCB = (float*)xmalloc(4 * (size_t)(rt + 1));
    float* y = (float*)xmalloc(4 * (size_t)n);
    float* xp = (float*)xmalloc(4 * (size_t)n);
    float* F = (float*)xmalloc(4 * (size_t)rmax * (size_t)rmax);
    float* fr = (float*)xmalloc(4 * (size_t)rmax);
    int rc = NR_OK;
    for (int32_t J = 0; J < ns; ++J) {
        const int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m, f = t->sn_first[J];
        for (int32_t e = 0; e < r * r; ++e) F[e] = 0.0f;
        for (int32_t e = 0; e < r; ++e) fr[e] = 0.0f;
        for (int32_t q = t->sn_a_ptr[J]; q < t->sn_a_ptr[J + 1]; ++q)
            F[t->sn_a_li[q] * r + t->sn_a_lj[q]] += a_vals[t->sn_a_slot[q]];
        for (int32_t k = 0; k < c; ++k) fr[k] += rhs[t->perm[f + k]];
        for (int32_t q = t->sn_child_ptr[J]; q < t->sn_child_ptr[J + 1]; ++q) {
            const int32_t C = t->sn_child[q], mc = t->sn_nrows[C];
            const int32_t* Rc = t->sn_rows + t->sn_row_off[C];
            int32_t* rel = (int32_t*)xmalloc(4 * (size_t)(mc + 1));
            for (int32_t a = 0; a < mc; ++a) rel[a] = front_local(t, J, Rc[a]);
            const float* cb = CB + cb_off[C];
            for (int32_t a = 0; a < mc; ++a)
                for (int32_t b = 0; b < mc; ++b) F[rel[a] * r + rel[b]] += cb[a * mc + b];
            for (int32_t a = 0; a < mc; ++a) fr[rel[a]] += RCB[rcb_off[C] + a];
            free(rel);
        }
        for (int32_t k = 0; k < c; ++k) {
            const float piv = F[k * r + k];
            if (!(fabsf(piv) >= 1e-25f)) rc = NR_STATUS_SINGULAR;
            for (int32_t i = k + 1; i < r; ++i) F[i * r + k] = F[i * r + k] / piv;
            for (int32_t i = k + 1; i < r; ++i) {
                const float l = F[i * r + k];
                for (int32_t j = k + 1; j < r; ++j) F[i * r + j] = fmaf(-l, F[k * r + j], F[i * r + j]);
                fr[i] = fmaf(-l, fr[k], fr[i]);
            }
        }
        memcpy(U + u_off[J], F, 4 * (size_t)c * (size_t)r);
        for (int32_t k = 0; k < c; ++k) y[f + k] = fr[k];
        for (int32_t a = 0; a < m; ++a) {
            for (int32_t b = 0; b < m; ++b) CB[cb_off[J] + a * m + b] = F[(c + a) * r + c + b];
            RCB[rcb_off[J] + a] = fr[c + a];
        }
    }
    if (rc == NR_OK) {
        for (int32_t J = ns - 1; J >= 0; --J) {
            const int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m, f = t->sn_first[J];
            const float* Uj = U + u_off[J];
            const int32_t* R = t->sn_rows + t->sn_row_off[J];
            for (int32_t k = 0; k < c; ++k) {
                float s = y[f + k];
                for (int32_t a = 0; a < m; ++a) s = fmaf(-Uj[k * r + c + a], xp[R[a]], s);
                fr[k] = s;
            }
            for (int32_t k = c - 1; k >= 0; --k) {
                const float xk = fr[k] / Uj[k * r + k];
                xp[f + k] = xk;
                for (int32_t i = 0; i < k; ++i) fr[i] = fmaf(-Uj[i * r + k], xk, fr[i]);
            }
        }
        for (int32_t i = 0; i < n; ++i) x[t->perm[i]] = xp[i];
    } else {
        for (int32_t i = 0; i < n; ++i) x[i] = 0.0f;
    }
    free(u_off); free(cb_off); free(rcb_off); free(U); free(CB); free(RCB); free(y); free(xp); free(F); free(fr);
    return rc;
}

/* The product's default route (NR_LU_AUTO).  The multifrontal LU does not
 * pivot; on circuits whose Jacobian needs row exchanges for stability (long
 * inverter chains, where a stage's gm exceeds its output conductance by the
 * stage gain) its element growth leaves a solution whose residual is many
 * orders above rounding level.  Such a solve is redone by the dense LU with
 * partial pivoting the reference runs (solveLU_Pivoted, math/linalg.h:
 * 742-854, reached from solvers/circuitsim_solver.cpp:291-295).  The test is
 * the relative residual max_i |b_i - (A x)_i| <= bound * max_i |b_i|, every
 * row's residual accumulated in CSR order, so that the HIP check reaches the
 * same decision.  (A normwise backward error, residual over |A| |x| + |b|,
 * would pass a float solve that has grown to 1e16 V: it is small whenever x
 * is large.) */
int orc_lu_flag(const orc_topo* t, const double* a_vals, const double* rhs, const double* x, double bound,
                double* eta) {
    const int32_t n = t->n;
    double mr = 0.0, mb = 0.0;
    int bad = 0;
    for (int32_t i = 0; i < n; ++i) {
        double r = rhs[i];
        for (int32_t k = t->row_ptr[i]; k < t->row_ptr[i + 1]; ++k) r = r - a_vals[k] * x[t->col_idx[k]];
        if (!isfinite(r) || !isfinite(x[i]) || !isfinite(rhs[i])) bad = 1;
        if (fabs(r) > mr) mr = fabs(r);
        if (fabs(rhs[i]) > mb) mb = fabs(rhs[i]);
    }
    if (eta) *eta = bad ? INFINITY : mr / mb;
    return bad || !(mr <= bound * mb);
}

int orc_lu_auto(const orc_topo* t, const double* a_vals, const double* rhs, double* x, int32_t* fallback) {
    if (fallback) *fallback = 0;
    int rc = orc_lu_multifrontal(t, a_vals, rhs, x, NULL);
    /* above NR_DENSE_MAX_N nodes the route is the multifrontal solve alone */
    if (t->n > NR_DENSE_MAX_N) return rc;
    if (rc == NR_OK && !orc_lu_flag(t, a_vals, rhs, x, ORC_LU_ETA_FP64, NULL)) return NR_OK;
    if (fallback) *fallback = 1;
    return orc_lu_dense(t->n, t->row_ptr, t->col_idx, a_vals, rhs, x, NULL);
}

static double kcl_soa(const nr_netlist* nl, const orc_soa* s, const double* v, double* res, const tran_ctx* tc);

int orc_jacobian_f32(const nr_netlist* nl, const orc_topo* t, const double* v, double gmin, float* a32) {
    orc_soa* s = orc_tensorize(nl);
    double* vals = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* lin = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    int rc = assemble_f32(nl, s, t, v, gmin, NULL, vals, lin, a32);
    free(vals); free(lin);
    orc_soa_free(s);
    return rc;
}

/* One Newton step of the FP32 path at v: Jacobian and LU in float, rhs =
 * the FP64 Newton residual rounded to float, vn = v + the float correction
 * (FP64 update).  The Newton residual b - A v of the reference's Norton
 * system is the KCL audit vector minus the gmin diagonal current.
 * With `promote` (the AUTO route) a float solve the residual check rejects
 * sets *fallback = 2 and leaves vn unset: the caller takes the iteration
 * again on the FP64 path (whose AUTO route has the dense fallback up to
 * NR_DENSE_MAX_N nodes) and stays there. */
static int newton_step_f32(const nr_netlist* nl, const orc_soa* s, const orc_topo* t, const double* v,
                           double gmin, const tran_ctx* tc, int32_t lu_kind, int promote, double* vals, double* lin,
                           float* a32, float* r32, float* x32, double* res, double* vn, int32_t* fallback) {
    const int32_t n = nl->num_nodes;
    *fallback = 0;
    int rc = assemble_f32(nl, s, t, v, gmin, tc, vals, lin, a32);
    if (rc != NR_OK) return rc;
    kcl_soa(nl, s, v, res, tc);
    for (int32_t i = 0; i < n; ++i) r32[i] = (float)(res[i] - gmin * v[i]);
    int dense = (lu_kind == ORC_LU_DENSE);
    if (!dense) {
        rc = orc_lu_multifrontal_f32(t, a32, r32, x32);
        for (int32_t i = 0; i < n; ++i) vn[i] = (double)x32[i];
        if (lu_kind == ORC_LU_AUTO && (t->n <= NR_DENSE_MAX_N || promote)) {
            double* ad = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
            double* rd = (double*)xmalloc(sizeof(double) * (size_t)n);
            for (int32_t k = 0; k < t->nnz_a; ++k) ad[k] = (double)a32[k];
            for (int32_t i = 0; i < n; ++i) rd[i] = (double)r32[i];
            dense = rc != NR_OK || orc_lu_flag(t, ad, rd, vn, ORC_LU_ETA_FP32, NULL);
            *fallback = dense;
            free(ad); free(rd);
            if (dense && promote) {
                *fallback = 2;
                return NR_OK;
            }
        }
    }
    if (dense) {
        /* the dense LU solves the Newton correction with the FP64 Jacobian
         * (the FP64 path's A) and the FP64 residual: rounding the Jacobian
         * to float alone moves the step of an ill-conditioned system */
        double* rhs64 = (double*)xmalloc(sizeof(double) * (size_t)n);
        rc = stamp_with_soa(nl, s, t, v, gmin, vals, rhs64, tc);
        if (rc == NR_OK) {
            orc_assemble(t, vals, lin);
            for (int32_t i = 0; i < n; ++i) rhs64[i] = res[i] - gmin * v[i];
            rc = orc_lu_dense(n, t->row_ptr, t->col_idx, lin, rhs64, vn, NULL);
        }
        free(rhs64);
    }
    for (int32_t i = 0; i < n; ++i) vn[i] = v[i] + vn[i];
    return rc;
}

/* one FP64 Newton system solved as `lu_kind` says */
static int solve_f64(const orc_topo* t, int32_t lu_kind, const double* av, const double* rhs, double* x,
                     int32_t* fallback) {
    *fallback = 0;
    switch (lu_kind) {
        case ORC_LU_DENSE: return orc_lu_dense(t->n, t->row_ptr, t->col_idx, av, rhs, x, NULL);
        case ORC_LU_SPARSE: return orc_lu_sparse(t, av, rhs, x, NULL);
        case ORC_LU_AUTO: return orc_lu_auto(t, av, rhs, x, fallback);
        default: return orc_lu_multifrontal(t, av, rhs, x, NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* a14: seedInitialGuess (solvers/circuitsim_solver.cpp:32-116)             */
/* ------------------------------------------------------------------------ */
void orc_seed(const nr_netlist* nl, double* v) {
    const int32_t N = nl->num_nodes;
    for (int32_t i = 0; i < N; ++i) v[i] = 0.0;
    for (int32_t i = 0; i < nl->n_vsrc; ++i) {
        int32_t p = nl->vsrc_pos[i];
        if (p > 0 && p <= N) v[p - 1] = nl->vsrc_volts[i];
    }
    for (int32_t i = 0; i < nl->n_diode; ++i) {
        int32_t a = nl->dio_anode[i];
        if (a > 0 && a <= N) v[a - 1] = dmax_ref(v[a - 1], 0.6);
    }
    double vSupply = 5.0;
    for (int32_t i = 0; i < nl->n_vsrc; ++i)
        if (nl->vsrc_volts[i] > vSupply) vSupply = nl->vsrc_volts[i];
    for (int32_t i = 0; i < nl->n_isrc; ++i) {
        int32_t p = nl->isrc_pos[i];
        if (p <= 0 || p > N) continue;
        if (v[p - 1] == 0.0) v[p - 1] = vSupply * 0.5;   /* no BJTs on this path */
    }
}

/* ------------------------------------------------------------------------ */
/* a18: pnjlim / compute_vcrit (physics/device_physics.h:132-179) and       */
/* applyPnJunctionLimits (solvers/circuitsim_solver.cpp:118-169)            */
/* ------------------------------------------------------------------------ */
double orc_pnjlim(double vnew, double vold, double vt, double vcrit) {
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
        vlim = dmax_ref(vnew, vold - 5.0 * vt);
    }
    return vlim;
}

double orc_vcrit(double vt, double is) { return vt * log(vt / (sqrt(2.0) * is)); }

static void apply_pn_limits_soa(const nr_netlist* nl, const orc_soa* s,
                                double* v_new, const double* v_old) {
    const int32_t N = nl->num_nodes;
    const double ts = nl->ambient_temp_K / 300.15;
    for (int32_t i = 0; i < s->nD; ++i) {
        double Vt = s->dVt[i], Is = s->dIs[i];
        Vt *= ts;
        Is *= pow(ts, 3.0) * exp(1.11 / (Vt / ts) - 1.11 / Vt);
        int32_t p = s->da[i] - 1, q = s->dc[i] - 1;
        if (p < 0 && q < 0) continue;
        if (p >= N || q >= N) continue;
        double vcrit = orc_vcrit(Vt, Is);
        double vjo = (p >= 0 ? v_old[p] : 0.0) - (q >= 0 ? v_old[q] : 0.0);
        double vjn = (p >= 0 ? v_new[p] : 0.0) - (q >= 0 ? v_new[q] : 0.0);
        double vl = orc_pnjlim(vjn, vjo, Vt, vcrit);
        double corr = vjn - vl;
        if (fabs(corr) > 1e-12) {
            if (p >= 0 && q >= 0) { v_new[p] -= corr * 0.5; v_new[q] += corr * 0.5; }
            else if (p >= 0) v_new[p] -= corr;
            else if (q >= 0) v_new[q] += corr;
        }
    }
}

void orc_apply_pn_limits(const nr_netlist* nl, const orc_topo* t,
                         double* v_new, const double* v_old) {
    (void)t;
    orc_soa* s = orc_tensorize(nl);
    apply_pn_limits_soa(nl, s, v_new, v_old);
    orc_soa_free(s);
}

/* ------------------------------------------------------------------------ */
/* a19: calculatePhysicalResiduals (physics/circuitsim.h:1250-1671), DC     */
/* terms, netlist order.  The MOSFET term uses the SoA parameters          */
/* and the generic-kernel current (same equations as the stamp).            */
/* ------------------------------------------------------------------------ */
static double diode_i_clamped(double vd, double Is, double N, double Vt) {
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

static double kcl_soa(const nr_netlist* nl, const orc_soa* s, const double* v,
                      double* res, const tran_ctx* tc) {
    const int32_t n = nl->num_nodes;
    for (int32_t i = 0; i < n; ++i) res[i] = 0.0;
#define ADDRES(node, val) do { int32_t nd_ = (node); if (nd_ > 0 && nd_ <= n) res[nd_ - 1] += (val); } while (0)
    for (int32_t i = 0; i < nl->n_res; ++i) {
        double ir = (volt(v, n, nl->res_n1[i]) - volt(v, n, nl->res_n2[i])) / nl->res_ohms[i];
        ADDRES(nl->res_n1[i], -ir); ADDRES(nl->res_n2[i], ir);
    }
    for (int32_t i = 0; i < nl->n_vsrc; ++i) {
        double V = (tc && tc->vsrc) ? tc->vsrc[i] : nl->vsrc_volts[i];
        double isrc = (V - (volt(v, n, nl->vsrc_pos[i]) - volt(v, n, nl->vsrc_neg[i]))) / 1e-3;
        ADDRES(nl->vsrc_pos[i], isrc); ADDRES(nl->vsrc_neg[i], -isrc);
    }
    for (int32_t i = 0; i < nl->n_isrc; ++i) {
        const double I = (tc && tc->isrc) ? tc->isrc[i] : nl->isrc_amps[i];
        ADDRES(nl->isrc_pos[i], -I); ADDRES(nl->isrc_neg[i], I);
    }
    /* diodes at the V_T the stamps use (V_T * T / 300.15): the reference's
     * audit takes the unscaled V_T (physics/circuitsim.h:1346-1347), so away
     * from 300.15 K its own audit can never pass; fixed as D7's MOSFET audit
     * is.  At 300.15 K the scale is exactly 1. */
    const double tscale = nl->ambient_temp_K / 300.15;
    for (int32_t i = 0; i < nl->n_diode; ++i) {
        double vd = volt(v, n, nl->dio_anode[i]) - volt(v, n, nl->dio_cathode[i]);
        double id = diode_i_clamped(vd, nl->dio_is[i], nl->dio_n[i], nl->dio_vt[i] * tscale);
        ADDRES(nl->dio_anode[i], -id); ADDRES(nl->dio_cathode[i], id);
    }
    if (s->nM > 0) {
        int32_t* pos = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)s->nM);
        for (int32_t p = 0; p < s->nM; ++p) pos[s->m_src[p]] = p;
        for (int32_t i = 0; i < nl->n_mos; ++i) {
            int32_t p = pos[i];
            double vgs, vds, ids, gm, gds;
            mos_one(volt(v, n, s->mg[p]), volt(v, n, s->md[p]), volt(v, n, s->ms[p]), s->mW[p], s->mL[p],
                    s->mKp[p], s->mVth[p], s->mLam[p], s->mP[p], &vgs, &vds, &ids, &gm, &gds);
            ADDRES(nl->mos_d[i], -ids); ADDRES(nl->mos_s[i], ids);
        }
        free(pos);
    }
    /* capacitor displacement current of the attempt's method
     * (physics/circuitsim.h:1460-1490) */
    for (int32_t i = 0; tc && i < nl->n_cap; ++i) {
        const double vc = volt(v, n, nl->cap_n1[i]) - volt(v, n, nl->cap_n2[i]);
        const double C = nl->cap_farads[i];
        double ic;
        if (tc->method == ORC_INT_TRAP) {
            const double geq = (2.0 * C) / tc->h;
            ic = geq * (vc - tc->hv0[i]) - tc->hi0[i];
        } else if (tc->method == ORC_INT_GEAR2) {
            ic = (C / tc->h) * (1.5 * vc - 2.0 * tc->hv0[i] + 0.5 * tc->hv1[i]);
        } else {
            ic = (C / tc->h) * (vc - tc->hv0[i]);
        }
        ADDRES(nl->cap_n1[i], -ic); ADDRES(nl->cap_n2[i], ic);
    }
#undef ADDRES
    double mx = 0.0;
    for (int32_t i = 0; i < n; ++i) mx = dmax_ref(mx, fabs(res[i]));
    return mx;
}

double orc_kcl(const nr_netlist* nl, const orc_topo* t, const double* v, double* per_node) {
    (void)t;
    orc_soa* s = orc_tensorize(nl);
    double* res = per_node ? per_node : (double*)xmalloc(sizeof(double) * (size_t)nl->num_nodes);
    double m = kcl_soa(nl, s, v, res, NULL);
    if (!per_node) free(res);
    orc_soa_free(s);
    return m;
}

/* a20: checkConvergence (physics/circuitsim.h:1218-1241) */
int orc_check_convergence(int32_t n, const double* v_new, const double* v_old,
                          double kcl, double kcl_tol) {
    if (kcl < 1e-9) return 1;
    for (int32_t i = 0; i < n; ++i) {
        double diff = fabs(v_new[i] - v_old[i]);
        double limit = 1e-3 * dmax_ref(fabs(v_new[i]), fabs(v_old[i])) + 1e-6;
        if (diff > limit) return 0;
    }
    if (kcl > kcl_tol) return 0;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* a15: solveDC (solvers/circuitsim_solver.cpp:176-555)                     */
/* ------------------------------------------------------------------------ */
void orc_dc_opts_default(orc_dc_opts* o) {
    o->max_iter = 50; o->gmin_steps = 19; o->lu_kind = ORC_LU_AUTO;
    o->start_step = 0; o->use_seed = 0; o->precision = ORC_PREC_FP64;
    o->trace_cap = 0; o->polish = 0; o->trace = NULL;
}

/* The NR convergence diagnostics of solveDC (solvers/circuitsim_solver.cpp:
 * 407-444, 489-551) kept incrementally: classifyFailure (solvers/
 * convergence_engine.h:26-87) reads the whole history, but only through
 * whether some max |dv| exceeds 100 V, the last ten residuals, the number of
 * residual direction changes and the history length, which are updated
 * record by record here. */
typedef struct {
    double last[10];       /* residual of record k at k % 10 */
    int64_t n;
    int32_t changes, big_dv, stagnation, divergence;
} dc_diag;

static void diag_record(dc_diag* d, double residual, double max_dv, double prev_final, int32_t it) {
    /* stagnation (:430-434): the step's iteration > 0, the residual within
     * 1e-15 of the last and above KCL_TOL (1e-12, physics_constants.h:76) */
    if (it > 0 && fabs(prev_final - residual) < 1e-15 && residual > 1e-12) d->stagnation = 1;
    if (max_dv > 100.0) d->big_dv = 1;
    d->last[d->n % 10] = residual;
    d->n++;
    if (d->n >= 3) {
        const double h1 = d->last[(d->n - 1) % 10], h2 = d->last[(d->n - 2) % 10], h3 = d->last[(d->n - 3) % 10];
        if (h1 > h2 && h2 > h3) d->divergence = 1;           /* :436-441 */
        const double pd = h2 - h3, cd = h1 - h2;                /* classifyFailure :63-69 */
        if ((pd > 0 && cd < 0) || (pd < 0 && cd > 0)) d->changes++;
    }
}

static int32_t diag_classify(const dc_diag* d, int singular) {
    if (singular) return ORC_FAIL_SINGULAR;
    if (d->n == 0) return ORC_FAIL_NONE;
    if (d->big_dv) return ORC_FAIL_OVERFLOW;
    if (d->n >= 5) {
        int up = 1;
        for (int64_t i = d->n - 4; i < d->n; ++i)
            if (d->last[i % 10] <= d->last[(i - 1) % 10]) { up = 0; break; }
        if (up) return ORC_FAIL_DIVERGENCE;
    }
    if (d->n >= 4 && (double)d->changes / (double)(d->n - 2) > 0.6) return ORC_FAIL_LIMIT_CYCLE;
    if (d->n >= 10) {
        const double initial = d->last[(d->n - 10) % 10], final_val = d->last[(d->n - 1) % 10];
        if (initial > 0.0 && final_val > 0.5 * initial) return ORC_FAIL_SLOW;
    }
    return ORC_FAIL_NONE;
}

/* the incremental classification fed one record at a time (checked against
 * the reference's classifyFailure on whole histories) */
int32_t orc_classify_history(const double* residual, const double* max_dv, int64_t n, int32_t singular) {
    dc_diag d;
    memset(&d, 0, sizeof(d));
    for (int64_t k = 0; k < n; ++k) diag_record(&d, residual[k], max_dv[k], k > 0 ? residual[k - 1] : 0.0, 0);
    return diag_classify(&d, singular);
}

int orc_solve_dc(const nr_netlist* nl, const orc_topo* t, const orc_dc_opts* opts,
                 double* v_io, orc_dc_out* out) {
    orc_dc_opts o;
    if (opts) o = *opts; else orc_dc_opts_default(&o);
    memset(out, 0, sizeof(*out));
    const int32_t n = nl->num_nodes;
    if (n <= 0) return NR_ERR_INVALID;
    orc_topo* own = NULL;
    if (!t) { int rc = orc_topology_build(nl, &own); if (rc) return rc; t = own; }
    orc_soa* s = orc_tensorize(nl);
    double* v = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* vn = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* rhs = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* res = (double*)xmalloc(sizeof(double) * (size_t)n);
    double* vals = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* av = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    int p32 = (o.precision == ORC_PREC_FP32_REFINE);
    const int promote = p32 && o.lu_kind == ORC_LU_AUTO;
    float* a32 = p32 ? (float*)xmalloc(sizeof(float) * (size_t)t->nnz_a) : NULL;
    float* r32 = p32 ? (float*)xmalloc(sizeof(float) * (size_t)n) : NULL;
    float* x32 = p32 ? (float*)xmalloc(sizeof(float) * (size_t)n) : NULL;
    if (o.use_seed) memcpy(v, v_io, sizeof(double) * (size_t)n);
    else orc_seed(nl, v);
    int converged = 0, singular = 0;
    int32_t total = 0, ref_total = 0, steps_done = 0, polished = 0;
    double final_kcl = 0.0, last_gmin = 1.0;
    dc_diag dg;
    memset(&dg, 0, sizeof(dg));
    for (int32_t k = o.start_step; k < o.gmin_steps; ++k) {
        double gmin = pow(10.0, -(3.0 + k * 0.5));
        if (gmin < 1e-12) gmin = 1e-12;
        last_gmin = gmin;
        int gconv = 0, sing_step = 0;
        int32_t it = 0;
        for (; it < o.max_iter; ++it) {
            int rc;
            int32_t fb = 0;
            rc = NR_OK;
            if (p32) {
                rc = newton_step_f32(nl, s, t, v, gmin, NULL, o.lu_kind, promote, vals, av, a32, r32, x32, res, vn,
                                     &fb);
                if (fb == 2) { p32 = 0; fb = 1; }   /* the rest of the solve on the FP64 path */
            }
            if (!p32) {
                int32_t fb64 = 0;
                stamp_with_soa(nl, s, t, v, gmin, vals, rhs, NULL);
                orc_assemble(t, vals, av);
                rc = solve_f64(t, o.lu_kind, av, rhs, vn, &fb64);
                fb += fb64;
            }
            out->lu_fallbacks += fb;
            if (rc != NR_OK) {
                final_kcl = 1e9;
                singular = 1; sing_step = 1;
                ref_total += it + 1;           /* the reference's count: here and below */
                break;
            }
            for (int32_t i = 0; i < n; ++i) if (isnan(vn[i]) || isinf(vn[i])) vn[i] = 0.0;
            apply_pn_limits_soa(nl, s, vn, v);
            double kcl = kcl_soa(nl, s, vn, res, NULL);
            if (it > 0 && kcl >= final_kcl && polished == 0) {   /* polish: full Newton steps */
                for (int32_t i = 0; i < n; ++i) vn[i] = v[i] + 0.5 * (vn[i] - v[i]);
                kcl = kcl_soa(nl, s, vn, res, NULL);
                if (kcl >= final_kcl) {
                    for (int32_t i = 0; i < n; ++i) vn[i] = v[i] + 0.5 * (vn[i] - v[i]);
                    kcl = kcl_soa(nl, s, vn, res, NULL);
                }
            }
            {
                /* ConvergenceStep (:407-425): damping 0.5 when the (damped)
                 * residual is still not below the last; worst node as the
                 * reference fills it, from a vector holding the residual at
                 * every node (so node 1) */
                double max_dv = 0.0;
                for (int32_t i = 0; i < n; ++i) {
                    const double d = fabs(vn[i] - v[i]);
                    if (d > max_dv) max_dv = d;
                }
                const double damping = (it > 0 && kcl >= final_kcl && polished == 0) ? 0.5 : 1.0;
                if (o.trace_cap > 0 && o.trace) {
                    orc_conv_step* r = &o.trace[dg.n % o.trace_cap];
                    r->residual = kcl; r->max_dv = max_dv; r->damping = damping;
                    r->gmin_step = k; r->iteration = it; r->worst_node = 1; r->pad = 0;
                }
                diag_record(&dg, kcl, max_dv, final_kcl, it);
            }
            final_kcl = kcl;
            double tol = (k == o.gmin_steps - 1) ? 1e-12 : 1.0;
            if (orc_check_convergence(n, vn, v, kcl, tol)) {
                memcpy(v, vn, sizeof(double) * (size_t)n);
                if (k == o.gmin_steps - 1 && polished < o.polish && it + 1 < o.max_iter) {
                    ++polished;      /* one more Newton iteration from the converged iterate */
                    continue;
                }
                gconv = 1;
                break;
            }
            memcpy(v, vn, sizeof(double) * (size_t)n);
        }
        steps_done++;
        int32_t true_it = gconv ? it + 1 : (sing_step ? it + 1 : it);
        if (k < 32) out->iters_per_step[k] = true_it;
        total += true_it;
        ref_total += it + 1;
        if (k == o.gmin_steps - 1 && gconv) { converged = 1; break; }
        if (!gconv && !(final_kcl < 1.0)) break;
    }
    double fk = kcl_soa(nl, s, v, res, NULL);
    {
        /* final per-node residuals: RMS, worst node and its voltage
         * (:489-508) */
        double sum_sq = 0.0, max_abs = 0.0;
        int32_t worst = 0;
        double worst_v = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            const double a = fabs(res[i]);
            sum_sq += a * a;
            if (a > max_abs) { max_abs = a; worst = i + 1; worst_v = v[i]; }
        }
        out->rms = sqrt(sum_sq / (double)n);
        out->worst_node = worst;
        out->worst_v = worst_v;
    }
    out->history = dg.n;
    out->failure = diag_classify(&dg, singular);
    out->stagnation = dg.stagnation;
    out->divergence = dg.divergence;
    out->gmin_used = last_gmin < 1e-3;
    out->iterations = total;
    out->ref_iterations = ref_total;
    out->status = converged ? NR_STATUS_CONVERGED : (singular ? NR_STATUS_SINGULAR : NR_STATUS_NOT_CONVERGED);
    out->gmin_steps_done = steps_done;
    out->kcl = fk;
    memcpy(v_io, v, sizeof(double) * (size_t)n);
    free(v); free(vn); free(rhs); free(res); free(vals); free(av); free(a32); free(r32); free(x32);
    orc_soa_free(s);
    if (own) orc_topology_free(own);
    return NR_OK;
}

int orc_solve_dc_batch(const nr_netlist* nl, const orc_topo* t, int32_t B,
                       const double* const* fields, const orc_dc_opts* opts,
                       int32_t threads, double* v_out, orc_dc_out* outs) {
    if (!t || B <= 0) return NR_ERR_INVALID;
    const int32_t n = nl->num_nodes;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int32_t b = 0; b < B; ++b) {
        nr_netlist inst = *nl;
        if (fields) {
            if (fields[NR_FIELD_RES_OHMS]) inst.res_ohms = fields[NR_FIELD_RES_OHMS] + (size_t)b * nl->n_res;
            if (fields[NR_FIELD_DIO_IS]) inst.dio_is = fields[NR_FIELD_DIO_IS] + (size_t)b * nl->n_diode;
            if (fields[NR_FIELD_MOS_KP]) inst.mos_kp = fields[NR_FIELD_MOS_KP] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_MOS_VTO]) inst.mos_vto = fields[NR_FIELD_MOS_VTO] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_ISRC_AMPS]) inst.isrc_amps = fields[NR_FIELD_ISRC_AMPS] + (size_t)b * nl->n_isrc;
            if (fields[NR_FIELD_VSRC_VOLTS]) inst.vsrc_volts = fields[NR_FIELD_VSRC_VOLTS] + (size_t)b * nl->n_vsrc;
            if (fields[NR_FIELD_CAP_FARADS]) inst.cap_farads = fields[NR_FIELD_CAP_FARADS] + (size_t)b * nl->n_cap;
            if (fields[NR_FIELD_TEMP_K]) inst.ambient_temp_K = fields[NR_FIELD_TEMP_K][b];
        }
        orc_dc_opts o;
        if (opts) o = *opts; else orc_dc_opts_default(&o);
        if (o.trace) o.trace += (size_t)b * (size_t)o.trace_cap;
        orc_solve_dc(&inst, t, &o, v_out + (size_t)b * n, &outs[b]);
    }
    return NR_OK;
}

/* ------------------------------------------------------------------------ */
/* a16: fixed-step backward-Euler transient, stepTransient restated          */
/* (solvers/circuitsim_solver.cpp:559-780)                                   */
/* ------------------------------------------------------------------------ */
/* PULSE value at time `time` (physics/circuitsim.h:2155-2170);
 * p = v1 v2 td tr tf pw per */
double orc_pulse_value(const double* p, double time) {
    double v1 = p[0], v2 = p[1], td = p[2], tr = p[3], tf = p[4], pw = p[5];
    double t = time - td;
    if (!(t > 0)) return v1;
    double per = p[6] > 0 ? p[6] : 1e99;
    t = fmod(t, per);
    if (t < tr) return v1 + (v2 - v1) * (t / tr);
    if (t < tr + pw) return v2;
    if (t < tr + pw + tf) return v2 + (v1 - v2) * ((t - tr - pw) / tf);
    return v1;
}

/* SINE value at `time` (physics/circuitsim.h:2171-2180);
 * p = vo va freq td theta phase_deg */
double orc_sine_value(const double* p, double time) {
    const double vo = p[0], va = p[1], freq = p[2], td = p[3], theta = p[4], ph = p[5];
    double t = time - td;
    if (t > 0) {
        double omega = 2.0 * 3.141592653589793 * freq;
        double damping = exp(-t * theta);
        double phase = ph * (3.141592653589793 / 180.0);
        return vo + va * sin(omega * t + phase) * damping;
    }
    return vo;
}

double orc_wave_value(int32_t kind, const double* par, double dc, double time) {
    if (kind == NR_WAVE_PULSE) return orc_pulse_value(par, time);
    if (kind == NR_WAVE_SINE) return orc_sine_value(par, time);
    return dc;
}

void orc_tran_opts_default(orc_tran_opts* o) {
    o->h = 1e-9; o->steps = 200; o->max_iter = 50; o->start = 0; o->lu_kind = ORC_LU_AUTO;
    o->precision = ORC_PREC_FP64; o->method = ORC_INT_GEAR1; o->control = ORC_TRAN_FIXED; o->t_stop = 0.0;
    o->wave_cap = 0; o->decimation = 1;
}

int32_t orc_wave_capacity(const orc_tran_opts* o) {
    if (o->wave_cap > 0) return o->wave_cap;
    if (o->control == ORC_TRAN_ARBITER) return (int32_t)ceil(o->t_stop / o->h) + 1;
    return o->steps + 1;
}

/* IntegratorArbiter (physics/circuitsim.h:57-368), one per instance. */
typedef struct {
    double *prev, *prev2, *dvdt;   /* [n]: last two stable solutions, their dV/dt */
    double prev_energy, prev_dt;
    int32_t instab, taa_init;
} orc_arbiter;

/* rejectHistory (:90-121): nodes whose quadratic prediction error
 * |d2V/dt2| dt^2 exceeds 10 mV get the linear extrapolation; returns how
 * many.  dV/dt_current and the stored dV/dt are the same quotient of the
 * same operands (analyzeStability stores (v - prev) / dt together with
 * prev2 = prev, prev = v, prev_dt = dt), so the acceleration is exactly 0
 * and nothing is ever rejected; the count checks that. */
static int32_t arb_reject_history(const orc_arbiter* a, int32_t n, double* pred, double dt) {
    if (!a->taa_init || dt <= 0 || a->prev_dt <= 0) return 0;
    int32_t cnt = 0;
    const double dt_sq = dt * dt;
    for (int32_t i = 0; i < n; ++i) {
        double dv_dt_current = (a->prev[i] - a->prev2[i]) / a->prev_dt;
        double ddv_dt2 = (dv_dt_current - a->dvdt[i]) / (0.5 * (dt + a->prev_dt));
        double accel_error = fabs(ddv_dt2) * dt_sq;
        if (accel_error > 0.01) { pred[i] = a->prev[i] + dv_dt_current * dt; ++cnt; }
    }
    return cnt;
}

static inline double lte_norm(double v) { return 1e-6 + fabs(v) * 1e-3; }

/* calculateLTE (:177-246): max over nodes, capacitor branches and MOSFET
 * gate drives of the linear-predictor error scaled by dt / (dt + prev_dt)
 * and normalised by 1e-6 + 1e-3 |v|. */
static double arb_lte(const nr_netlist* nl, const orc_arbiter* a, int32_t n, const double* cur,
                      const double* hv0, const double* hv1, double dt) {
    if (dt <= 0 || a->prev_dt <= 0) return 0.0;
    const double c_factor = dt / (dt + a->prev_dt);
    double max_err = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        double v_pred = a->prev[i] + (a->prev[i] - a->prev2[i]) * (dt / a->prev_dt);
        double err = fabs(cur[i] - v_pred) * c_factor;
        max_err = dmax_ref(max_err, err / lte_norm(cur[i]));
    }
    for (int32_t i = 0; i < nl->n_cap; ++i) {
        const int32_t p = nl->cap_n1[i], q = nl->cap_n2[i];
        double v1 = (p > 0) ? cur[p - 1] : 0.0;
        double v2 = (q > 0) ? cur[q - 1] : 0.0;
        double v_branch = v1 - v2;
        double v_pred = hv0[i] + (hv0[i] - hv1[i]) * (dt / a->prev_dt);
        double err = fabs(v_branch - v_pred) * c_factor;
        max_err = dmax_ref(max_err, err / lte_norm(v_branch));
    }
    for (int32_t i = 0; i < nl->n_mos; ++i) {
        const int32_t g = nl->mos_g[i], so = nl->mos_s[i];
        double vgs = volt(cur, n, g) - volt(cur, n, so);
        double vgs_p = volt(a->prev, n, g) - volt(a->prev, n, so);
        double vgs_pp = volt(a->prev2, n, g) - volt(a->prev2, n, so);
        double vgs_pred = vgs_p + (vgs_p - vgs_pp) * (dt / a->prev_dt);
        double err = fabs(vgs - vgs_pred) * c_factor;
        max_err = dmax_ref(max_err, err / lte_norm(vgs));
    }
    return max_err;
}

/* stored capacitor energy sum 0.5 C v^2, netlist order (:289-295) */
static double cap_energy(const nr_netlist* nl, int32_t n, const double* v) {
    double e = 0.0;
    for (int32_t i = 0; i < nl->n_cap; ++i) {
        double x = 0.0;
        if (nl->cap_n1[i] > 0 && nl->cap_n1[i] <= n) x += v[nl->cap_n1[i] - 1];
        if (nl->cap_n2[i] > 0 && nl->cap_n2[i] <= n) x -= v[nl->cap_n2[i] - 1];
        e += 0.5 * nl->cap_farads[i] * x * x;
    }
    return e;
}

/* analyzeStability (:276-367): 0 stable (the arbiter takes the solution),
 * 1 reduce the step, 2 switch the method.  The source power uses the DC
 * current value, as the reference does. */
static int arb_analyze(const nr_netlist* nl, orc_arbiter* a, int32_t n, const double* cur, const double* hv0,
                       const double* hv1, double dt) {
    double lte = arb_lte(nl, a, n, cur, hv0, hv1, dt);
    int high_lte = (lte > 1.0);
    double current_energy = cap_energy(nl, n, cur);
    double power_in = 0.0;
    for (int32_t i = 0; i < nl->n_isrc; ++i) {
        double v_diff = 0.0;
        if (nl->isrc_pos[i] > 0 && nl->isrc_pos[i] <= n) v_diff += cur[nl->isrc_pos[i] - 1];
        if (nl->isrc_neg[i] > 0 && nl->isrc_neg[i] <= n) v_diff -= cur[nl->isrc_neg[i] - 1];
        power_in += fabs(nl->isrc_amps[i] * v_diff);
    }
    double energy_tolerance = 1e-9 + power_in * dt * 1.5;
    int exploding = (current_energy > a->prev_energy + energy_tolerance + 1e-6) &&
                    (current_energy > a->prev_energy * 1.5);
    if (high_lte || exploding) {
        a->instab++;
        return a->instab >= 3 ? 2 : 1;
    }
    a->instab = a->instab - 1 > 0 ? a->instab - 1 : 0;
    if (a->prev_dt > 0) {
        for (int32_t i = 0; i < n; ++i) a->dvdt[i] = (cur[i] - a->prev[i]) / dt;
        a->taa_init = 1;
    }
    memcpy(a->prev2, a->prev, sizeof(double) * (size_t)n);
    memcpy(a->prev, cur, sizeof(double) * (size_t)n);
    a->prev_energy = current_energy;
    a->prev_dt = dt;
    return 0;
}

/* analyzeStability on a given state (for tests): prev / prev2 [n],
 * scalars in/out; returns the suggestion, *lte_out the LTE */
int orc_arbiter_analyze(const nr_netlist* nl, const double* cur, const double* hv0, const double* hv1, double dt,
                        double* prev, double* prev2, double* prev_energy, double* prev_dt, int32_t* instab,
                        double* lte_out) {
    const int32_t n = nl->num_nodes;
    orc_arbiter a;
    a.prev = prev; a.prev2 = prev2;
    a.dvdt = (double*)xmalloc(sizeof(double) * (size_t)n);
    a.prev_energy = *prev_energy; a.prev_dt = *prev_dt; a.instab = *instab; a.taa_init = 0;
    if (lte_out) *lte_out = arb_lte(nl, &a, n, cur, hv0, hv1, dt);
    const int r = arb_analyze(nl, &a, n, cur, hv0, hv1, dt);
    *prev_energy = a.prev_energy; *prev_dt = a.prev_dt; *instab = a.instab;
    free(a.dvdt);
    return r;
}

/* Transient of one instance: the restated stepTransient
 * (solvers/circuitsim_solver.cpp:559-780) driven as its caller drives it
 * (engine_impl.cpp:205-256: every step asks for h; the loop runs while
 * t < t_stop), or `steps` fixed steps of h. */
int orc_solve_transient(const nr_netlist* nl, const orc_topo* t, const orc_tran_opts* opts, double* v_io,
                        int32_t n_probe, const int32_t* probe, double* wave_t, double* wave, orc_tran_out* out) {
    orc_tran_opts o;
    if (opts) o = *opts; else orc_tran_opts_default(&o);
    memset(out, 0, sizeof(*out));
    const int32_t n = nl->num_nodes;
    const int fixed = (o.control == ORC_TRAN_FIXED);
    if (n <= 0 || !t || !t->transient || !(o.h >= 1e-18) || o.decimation < 1) return NR_ERR_INVALID;
    if (fixed ? o.steps <= 0 : !(o.t_stop > 0)) return NR_ERR_INVALID;
    if (o.method < ORC_INT_TRAP || o.method > ORC_INT_GEAR2) return NR_ERR_INVALID;
    const int32_t cap = orc_wave_capacity(&o);
    orc_soa* s = orc_tensorize(nl);
    const size_t nb = sizeof(double) * (size_t)n;
    double* v = (double*)xmalloc(nb);
    double* vn = (double*)xmalloc(nb);
    double* rhs = (double*)xmalloc(nb);
    double* res = (double*)xmalloc(nb);
    double* vals = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* av = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    double* vs = (double*)xmalloc(sizeof(double) * (size_t)(nl->n_vsrc + 1));
    double* is = (double*)xmalloc(sizeof(double) * (size_t)(nl->n_isrc + 1));
    double* hv0 = (double*)xmalloc(sizeof(double) * (size_t)(nl->n_cap + 1));
    double* hv1 = (double*)xmalloc(sizeof(double) * (size_t)(nl->n_cap + 1));
    double* hi0 = (double*)xmalloc(sizeof(double) * (size_t)(nl->n_cap + 1));
    orc_arbiter arb;
    arb.prev = (double*)xmalloc(nb); arb.prev2 = (double*)xmalloc(nb); arb.dvdt = (double*)xmalloc(nb);
    int p32 = (o.precision == ORC_PREC_FP32_REFINE);
    const int promote = p32 && o.lu_kind == ORC_LU_AUTO;
    float* a32 = p32 ? (float*)xmalloc(sizeof(float) * (size_t)t->nnz_a) : NULL;
    float* r32 = p32 ? (float*)xmalloc(sizeof(float) * (size_t)n) : NULL;
    float* x32 = p32 ? (float*)xmalloc(sizeof(float) * (size_t)n) : NULL;
    /* start state: zero voltages, history and arbiter (the reference's
     * IntegratorArbiter::initialize, :135-173, and zero-filled history), or
     * the caller's state with the history and arbiter at it */
    if (o.start) memcpy(v, v_io, nb);
    else memset(v, 0, nb);
    for (int32_t i = 0; i < nl->n_cap; ++i) {
        hv0[i] = volt(v, n, nl->cap_n1[i]) - volt(v, n, nl->cap_n2[i]);
        hv1[i] = hv0[i];
        hi0[i] = 0.0;
    }
    memcpy(arb.prev, v, nb); memcpy(arb.prev2, v, nb); memset(arb.dvdt, 0, nb);
    arb.prev_energy = o.start ? cap_energy(nl, n, v) : 0.0;
    arb.prev_dt = 1e-9;
    arb.instab = 0; arb.taa_init = 0;
    if (wave_t) for (int32_t k = 0; k < cap; ++k) wave_t[k] = NAN;
    if (wave) for (int64_t k = 0; k < (int64_t)cap * n_probe; ++k) wave[k] = NAN;
    int64_t rec = 0;
    int32_t dec = 0;
#define RECORD(time_)                                                                       \
    do {                                                                                    \
        const int32_t slot_ = (int32_t)(rec % cap);                                         \
        if (wave_t) wave_t[slot_] = (time_);                                                \
        if (wave) for (int32_t p_ = 0; p_ < n_probe; ++p_)                                  \
            wave[(size_t)slot_ * n_probe + p_] = volt(v, n, probe[p_]);                     \
        ++rec;                                                                              \
    } while (0)
    RECORD(0.0);
    int32_t method = o.method, total = 0, steps_done = 0, rejected = 0, switched_at = -1, taa = 0;
    int32_t status = NR_STATUS_CONVERGED;
    double kcl = 0.0, tcur = 0.0, h_last = o.h;
    while (fixed ? steps_done < o.steps : tcur < o.t_stop) {
        double h = o.h;
        int32_t attempt = 0, fail_status = NR_STATUS_NOT_CONVERGED;
        int accepted = 0;
        while (attempt < (fixed ? 1 : 4)) {
            h_last = h;
            /* Newton estimate: the arbiter's last voltages (getPrevVoltages +
             * rejectHistory, :587-595); the fixed control continues from the
             * last accepted step */
            if (!fixed) {
                memcpy(v, arb.prev, nb);
                taa += arb_reject_history(&arb, n, v, h);
            }
            const double tn = tcur + h;
            for (int32_t i = 0; i < nl->n_vsrc; ++i)
                vs[i] = orc_wave_value(nl->vsrc_wave ? nl->vsrc_wave[i] : 0,
                                       nl->vsrc_wave_par ? nl->vsrc_wave_par + 7 * (size_t)i : NULL,
                                       nl->vsrc_volts[i], tn);
            for (int32_t i = 0; i < nl->n_isrc; ++i)
                is[i] = orc_wave_value(nl->isrc_wave ? nl->isrc_wave[i] : 0,
                                       nl->isrc_wave_par ? nl->isrc_wave_par + 7 * (size_t)i : NULL,
                                       nl->isrc_amps[i], tn);
            tran_ctx tc = {h, method, vs, is, hv0, hv1, hi0, 0};
            int conv = 0, sing = 0;
            int32_t it = 0;
            for (; it < o.max_iter; ++it) {
                int rc = NR_OK;
                int32_t fb = 0;
                if (p32) {
                    rc = newton_step_f32(nl, s, t, v, 1e-12, &tc, o.lu_kind, promote, vals, av, a32, r32, x32, res,
                                         vn, &fb);
                    if (fb == 2) { p32 = 0; fb = 1; }   /* the rest of the transient on the FP64 path */
                }
                if (!p32) {
                    int32_t fb64 = 0;
                    stamp_with_soa(nl, s, t, v, 1e-12, vals, rhs, &tc);
                    orc_assemble(t, vals, av);
                    rc = solve_f64(t, o.lu_kind, av, rhs, vn, &fb64);
                    fb += fb64;
                }
                out->lu_fallbacks += fb;
                if (rc != NR_OK) { sing = 1; break; }
                for (int32_t i = 0; i < n; ++i) if (isnan(vn[i]) || isinf(vn[i])) vn[i] = 0.0;
                kcl = kcl_soa(nl, s, vn, res, &tc);
                int c = orc_check_convergence(n, vn, v, kcl, 1e-12);
                memcpy(v, vn, nb);
                if (c) { conv = 1; break; }
            }
            if (sing) { total += it + 1; kcl = 1e9; fail_status = NR_STATUS_SINGULAR; }
            else if (!conv) { total += o.max_iter; fail_status = NR_STATUS_NOT_CONVERGED; }
            else total += it + 1;
            if (!conv) { ++rejected; h *= 0.5; ++attempt; continue; }
            if (fixed) { accepted = 1; break; }
            const int sug = arb_analyze(nl, &arb, n, v, hv0, hv1, h);
            if (sug == 1) { ++rejected; h *= 0.5; ++attempt; fail_status = NR_STATUS_NOT_CONVERGED; continue; }
            if (sug == 2 && method == ORC_INT_TRAP) {
                ++rejected;
                method = ORC_INT_GEAR2;
                switched_at = steps_done;
                continue;
            }
            accepted = 1;
            break;
        }
        if (!accepted) { status = fail_status; break; }
        /* history of the accepted step, with the method in effect
         * (solvers/circuitsim_solver.cpp:668-699) */
        for (int32_t i = 0; i < nl->n_cap; ++i) {
            double vc = 0.0;
            if (nl->cap_n1[i] > 0 && nl->cap_n1[i] <= n) vc += v[nl->cap_n1[i] - 1];
            if (nl->cap_n2[i] > 0 && nl->cap_n2[i] <= n) vc -= v[nl->cap_n2[i] - 1];
            const double C = nl->cap_farads[i];
            double ic;
            if (method == ORC_INT_TRAP) ic = (2.0 * C / h) * (vc - hv0[i]) - hi0[i];
            else if (method == ORC_INT_GEAR1) ic = (C / h) * (vc - hv0[i]);
            else if (hv1[i] == 0.0 && tcur > h) ic = (C / h) * (vc - hv0[i]);
            else ic = (C / h) * (1.5 * vc - 2.0 * hv0[i] + 0.5 * hv1[i]);
            hv1[i] = hv0[i];
            hv0[i] = vc;
            hi0[i] = ic;
        }
        tcur = tcur + h;
        ++steps_done;
        if (++dec >= o.decimation) { dec = 0; RECORD(tcur); }
    }
#undef RECORD
    out->iterations = total;
    out->status = status;
    out->steps_done = steps_done;
    out->rejected = rejected;
    out->method = method;
    out->switched_at = switched_at;
    out->instability = arb.instab;
    out->taa_rejections = taa;
    out->recorded = rec;
    out->kcl = kcl;
    out->t_end = tcur;
    out->h_last = h_last;
    memcpy(v_io, v, nb);
    free(v); free(vn); free(rhs); free(res); free(vals); free(av); free(vs); free(is);
    free(hv0); free(hv1); free(hi0); free(arb.prev); free(arb.prev2); free(arb.dvdt);
    free(a32); free(r32); free(x32);
    orc_soa_free(s);
    return NR_OK;
}

int orc_solve_transient_batch(const nr_netlist* nl, const orc_topo* t, int32_t B, const double* const* fields,
                              const orc_tran_opts* opts, int32_t threads, double* v_io, int32_t n_probe,
                              const int32_t* probe, double* wave_t, double* wave, orc_tran_out* outs) {
    if (!t || B <= 0) return NR_ERR_INVALID;
    const int32_t n = nl->num_nodes;
    orc_tran_opts o;
    if (opts) o = *opts; else orc_tran_opts_default(&o);
    const int32_t cap = orc_wave_capacity(&o);
    int bad = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
    for (int32_t b = 0; b < B; ++b) {
        nr_netlist inst = *nl;
        if (fields) {
            if (fields[NR_FIELD_RES_OHMS]) inst.res_ohms = fields[NR_FIELD_RES_OHMS] + (size_t)b * nl->n_res;
            if (fields[NR_FIELD_DIO_IS]) inst.dio_is = fields[NR_FIELD_DIO_IS] + (size_t)b * nl->n_diode;
            if (fields[NR_FIELD_MOS_KP]) inst.mos_kp = fields[NR_FIELD_MOS_KP] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_MOS_VTO]) inst.mos_vto = fields[NR_FIELD_MOS_VTO] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_ISRC_AMPS]) inst.isrc_amps = fields[NR_FIELD_ISRC_AMPS] + (size_t)b * nl->n_isrc;
            if (fields[NR_FIELD_VSRC_VOLTS]) inst.vsrc_volts = fields[NR_FIELD_VSRC_VOLTS] + (size_t)b * nl->n_vsrc;
            if (fields[NR_FIELD_CAP_FARADS]) inst.cap_farads = fields[NR_FIELD_CAP_FARADS] + (size_t)b * nl->n_cap;
            if (fields[NR_FIELD_TEMP_K]) inst.ambient_temp_K = fields[NR_FIELD_TEMP_K][b];
        }
        bad |= orc_solve_transient(&inst, t, &o, v_io + (size_t)b * n, n_probe, probe,
                                   wave_t ? wave_t + (size_t)b * cap : NULL,
                                   wave ? wave + (size_t)b * cap * n_probe : NULL, &outs[b]) != NR_OK;
    }
    return bad ? NR_ERR_INVALID : NR_OK;
}

/* SoA arrays of tensorizeNetlist (same layout as the reference harness's
 * ref_tensorize, for the fixture comparison). */
int orc_tensorize_export(const nr_netlist* nl, int32_t* d_a, int32_t* d_c, double* d_is, double* d_n,
                         double* d_vt, int32_t* m_d, int32_t* m_g, int32_t* m_s, int32_t* m_b,
                         double* m_w, double* m_l, double* m_kp, double* m_vth, double* m_lam,
                         int32_t* m_pmos) {
    orc_soa* s = orc_tensorize(nl);
    for (int32_t i = 0; i < s->nD; ++i) {
        d_a[i] = s->da[i]; d_c[i] = s->dc[i]; d_is[i] = s->dIs[i]; d_n[i] = s->dN[i]; d_vt[i] = s->dVt[i];
    }
    for (int32_t i = 0; i < s->nM; ++i) {
        m_d[i] = s->md[i]; m_g[i] = s->mg[i]; m_s[i] = s->ms[i]; m_b[i] = s->mb[i];
        m_w[i] = s->mW[i]; m_l[i] = s->mL[i]; m_kp[i] = s->mKp[i]; m_vth[i] = s->mVth[i];
        m_lam[i] = s->mLam[i]; m_pmos[i] = s->mP[i];
    }
    orc_soa_free(s);
    return 0;
}

/* diode_model_pair current used by the KCL audit */
double orc_diode_i_clamped(double vd, double Is, double N, double Vt) { return diode_i_clamped(vd, Is, N, Vt); }

/* ---- AC small-signal analysis (ACSolver::solveAC, solvers/ac_solver.h:339-383) ----
 * Complex arithmetic with explicit rounding steps, the formulas the GPU's
 * complex path uses:
 *   a + b = (ar + br, ai + bi)
 *   fma(a, b, c) = (fma(ar, br, fma(-ai, bi, cr)), fma(ar, bi, fma(ai, br, ci)))
 *   a / b = (fma(ar, br, ai bi) / d, fma(ai, br, -(ar bi)) / d), d = fma(br, br, bi bi)
 *   a pivot is zero when max(|re|, |im|) < 1e-25 */
typedef struct { double re, im; } cplx;
static inline cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cplx c_neg(cplx a) { cplx r = {-a.re, -a.im}; return r; }
static inline cplx c_fma(cplx a, cplx b, cplx c) {
    cplx r = {fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im))};
    return r;
}
static inline cplx c_div(cplx a, cplx b) {
    const double d = fma(b.re, b.re, b.im * b.im);
    cplx r = {fma(a.re, b.re, a.im * b.im) / d, fma(a.im, b.re, -(a.re * b.im)) / d};
    return r;
}
static inline double c_mag(cplx a) {
    const double x = fabs(a.re), y = fabs(a.im);
    return x >= y ? x : y;
}

/* orc_lu_multifrontal on complex values (same operation order).  With
 * transpose = 1 it factors and solves A^T instead: each A entry (i, j) is
 * placed at the front position of (j, i), which lies in the same front
 * because the filled pattern is symmetric. */
static int lu_multifrontal_c(const orc_topo* t, const cplx* a_vals, const cplx* rhs, cplx* x, int transpose) {
    const int32_t n = t->n, ns = t->ns;
    int64_t* u_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t* cb_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t* rcb_off = (int64_t*)xmalloc(8 * (size_t)ns);
    int64_t ut = 0, ct = 0, rt = 0;
    int32_t rmax = 1;
    for (int32_t J = 0; J < ns; ++J) {
        int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m;
        u_off[J] = ut; ut += (int64_t)c * r;
        cb_off[J] = ct; ct += (int64_t)m * m;
        rcb_off[J] = rt; rt += m;
        if (r > rmax) rmax = r;
    }
    const size_t cz = sizeof(cplx);
    cplx* U = (cplx*)xmalloc(cz * (size_t)(ut + 1));
    cplx* CB = (cplx*)xmalloc(cz * (size_t)(ct + 1));
    cplx* RCB = (cplx*)xmalloc(cz * (size_t)(rt + 1));
    cplx* y = (cplx*)xmalloc(cz * (size_t)n);
    cplx* xp = (cplx*)xmalloc(cz * (size_t)n);
    cplx* F = (cplx*)xmalloc(cz * (size_t)rmax * (size_t)rmax);
    cplx* fr = (cplx*)xmalloc(cz * (size_t)rmax);
    const cplx zero = {0.0, 0.0};
    int rc = NR_OK;
    for (int32_t J = 0; J < ns; ++J) {
        const int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m, f = t->sn_first[J];
        for (int32_t e = 0; e < r * r; ++e) F[e] = zero;
        for (int32_t e = 0; e < r; ++e) fr[e] = zero;
        for (int32_t q = t->sn_a_ptr[J]; q < t->sn_a_ptr[J + 1]; ++q) {
            cplx* d = transpose ? &F[t->sn_a_lj[q] * r + t->sn_a_li[q]] : &F[t->sn_a_li[q] * r + t->sn_a_lj[q]];
            *d = c_add(*d, a_vals[t->sn_a_slot[q]]);
        }
        for (int32_t k = 0; k < c; ++k) fr[k] = c_add(fr[k], rhs[t->perm[f + k]]);
        for (int32_t q = t->sn_child_ptr[J]; q < t->sn_child_ptr[J + 1]; ++q) {
            const int32_t C = t->sn_child[q], mc = t->sn_nrows[C];
            const int32_t* Rc = t->sn_rows + t->sn_row_off[C];
            int32_t* rel = (int32_t*)xmalloc(4 * (size_t)(mc + 1));
            for (int32_t a = 0; a < mc; ++a) rel[a] = front_local(t, J, Rc[a]);
            const cplx* cb = CB + cb_off[C];
            for (int32_t a = 0; a < mc; ++a)
                for (int32_t b = 0; b < mc; ++b) F[rel[a] * r + rel[b]] = c_add(F[rel[a] * r + rel[b]], cb[a * mc + b]);
            for (int32_t a = 0; a < mc; ++a) fr[rel[a]] = c_add(fr[rel[a]], RCB[rcb_off[C] + a]);
            free(rel);
        }
        for (int32_t k = 0; k < c; ++k) {
            const cplx piv = F[k * r + k];
            if (!(c_mag(piv) >= 1e-25)) rc = NR_STATUS_SINGULAR;
            for (int32_t i = k + 1; i < r; ++i) F[i * r + k] = c_div(F[i * r + k], piv);
            for (int32_t i = k + 1; i < r; ++i) {
                const cplx l = c_neg(F[i * r + k]);
                for (int32_t j = k + 1; j < r; ++j) F[i * r + j] = c_fma(l, F[k * r + j], F[i * r + j]);
                fr[i] = c_fma(l, fr[k], fr[i]);
            }
        }
        memcpy(U + u_off[J], F, cz * (size_t)c * (size_t)r);
        for (int32_t k = 0; k < c; ++k) y[f + k] = fr[k];
        for (int32_t a = 0; a < m; ++a) {
            for (int32_t b = 0; b < m; ++b) CB[cb_off[J] + a * m + b] = F[(c + a) * r + c + b];
            RCB[rcb_off[J] + a] = fr[c + a];
        }
    }
    if (rc == NR_OK) {
        for (int32_t J = ns - 1; J >= 0; --J) {
            const int32_t c = t->sn_ncols[J], m = t->sn_nrows[J], r = c + m, f = t->sn_first[J];
            const cplx* Uj = U + u_off[J];
            const int32_t* R = t->sn_rows + t->sn_row_off[J];
            for (int32_t k = 0; k < c; ++k) {
                cplx s = y[f + k];
                for (int32_t a = 0; a < m; ++a) s = c_fma(c_neg(Uj[k * r + c + a]), xp[R[a]], s);
                fr[k] = s;
            }
            for (int32_t k = c - 1; k >= 0; --k) {
                const cplx xk = c_div(fr[k], Uj[k * r + k]);
                xp[f + k] = xk;
                for (int32_t i = 0; i < k; ++i) fr[i] = c_fma(c_neg(Uj[i * r + k]), xk, fr[i]);
            }
        }
        for (int32_t i = 0; i < n; ++i) x[t->perm[i]] = xp[i];
    } else {
        for (int32_t i = 0; i < n; ++i) x[i] = zero;
    }
    free(u_off); free(cb_off); free(rcb_off); free(U); free(CB); free(RCB); free(y); free(xp); free(F); free(fr);
    return rc;
}

/* log-spaced frequencies as solveAC spaces them (:353-359) */
int orc_ac_freqs(const orc_ac_opts* o, double* freqs) {
    if (o->n_points <= 0) return NR_ERR_INVALID;
    const double logStart = log10(dmax_ref(o->f_start, 1e-3));
    const double logStop = log10(dmax_ref(o->f_stop, o->f_start + 1.0));
    const int32_t den = o->n_points - 1 > 1 ? o->n_points - 1 : 1;
    for (int32_t i = 0; i < o->n_points; ++i) {
        const double logF = logStart + (logStop - logStart) * i / den;
        freqs[i] = pow(10.0, logF);
    }
    return NR_OK;
}

/* AC sweep of one instance at its operating point v_op: Y(jw) = G + jwC
 * stamped in the transient pattern's emission order (real and imaginary parts
 * alike), the 1e-12 diagonal, then the 1e6 S excitation at the input node
 * with 1e6 x amplitude injected (:364-375); out[f][p][2]. */
int orc_solve_ac(const nr_netlist* nl, const orc_topo* t, const double* v_op, const orc_ac_opts* o,
                 int32_t n_probe, const int32_t* probe, double* out, int32_t* status) {
    const int32_t n = nl->num_nodes;
    if (!t || !t->transient || o->n_points <= 0 || o->input_node < 0 || o->input_node > n) return NR_ERR_INVALID;
    orc_soa* s = orc_tensorize(nl);
    const int32_t F = o->n_points;
    double* freqs = (double*)xmalloc(sizeof(double) * (size_t)F);
    orc_ac_freqs(o, freqs);
    double* vre = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* vim = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* are = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    double* aim = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    cplx* A = (cplx*)xmalloc(sizeof(cplx) * (size_t)t->nnz_a);
    cplx* rhs = (cplx*)xmalloc(sizeof(cplx) * (size_t)n);
    cplx* x = (cplx*)xmalloc(sizeof(cplx) * (size_t)n);
    int32_t in_slot = -1;
    if (o->input_node > 0)
        for (int32_t k = t->row_ptr[o->input_node - 1]; k < t->row_ptr[o->input_node]; ++k)
            if (t->col_idx[k] == o->input_node - 1) in_slot = k;
    tran_ctx tc = {1.0, ORC_INT_GEAR1, NULL, NULL, NULL, NULL, NULL, 0};
    emitter e;
    memset(&e, 0, sizeof(e));
    e.vals = vre;
    e.ac = 1;
    stamp_all(nl, s, v_op, 1e-12, &e, &tc);
    int rc = e.nv == t->n_stamps ? NR_OK : NR_ERR_TOPOLOGY;
    orc_assemble(t, vre, are);
    *status = NR_STATUS_CONVERGED;
    for (int32_t f = 0; f < F && rc == NR_OK; ++f) {
        const double omega = 2.0 * 3.14159265358979323846 * freqs[f];
        memset(&e, 0, sizeof(e));
        e.vals = vim;
        e.ac = 2;
        e.omega = omega;
        stamp_all(nl, s, v_op, 1e-12, &e, &tc);
        orc_assemble(t, vim, aim);
        for (int32_t k = 0; k < t->nnz_a; ++k) { A[k].re = are[k]; A[k].im = aim[k]; }
        if (in_slot >= 0) { A[in_slot].re = A[in_slot].re + 1e6; A[in_slot].im = A[in_slot].im + 0.0; }
        for (int32_t i = 0; i < n; ++i) { rhs[i].re = 0.0; rhs[i].im = 0.0; }
        if (o->input_node > 0) {
            rhs[o->input_node - 1].re = 0.0 + 1e6 * o->amplitude;
            rhs[o->input_node - 1].im = 0.0 + 0.0;
        }
        if (lu_multifrontal_c(t, A, rhs, x, 0) != NR_OK) *status = NR_STATUS_SINGULAR;
        for (int32_t p = 0; p < n_probe; ++p) {
            out[((size_t)f * n_probe + p) * 2] = x[probe[p] - 1].re;
            out[((size_t)f * n_probe + p) * 2 + 1] = x[probe[p] - 1].im;
        }
    }
    free(freqs); free(vre); free(vim); free(are); free(aim); free(A); free(rhs); free(x);
    orc_soa_free(s);
    return rc;
}

int orc_solve_ac_batch(const nr_netlist* nl, const orc_topo* t, int32_t B, const double* const* fields,
                       const double* v_op, const orc_ac_opts* o, int32_t threads, int32_t n_probe,
                       const int32_t* probe, double* out, int32_t* status) {
    if (!t || B <= 0) return NR_ERR_INVALID;
    const int32_t n = nl->num_nodes;
    int bad = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
    for (int32_t b = 0; b < B; ++b) {
        nr_netlist inst = *nl;
        if (fields) {
            if (fields[NR_FIELD_RES_OHMS]) inst.res_ohms = fields[NR_FIELD_RES_OHMS] + (size_t)b * nl->n_res;
            if (fields[NR_FIELD_DIO_IS]) inst.dio_is = fields[NR_FIELD_DIO_IS] + (size_t)b * nl->n_diode;
            if (fields[NR_FIELD_MOS_KP]) inst.mos_kp = fields[NR_FIELD_MOS_KP] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_MOS_VTO]) inst.mos_vto = fields[NR_FIELD_MOS_VTO] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_ISRC_AMPS]) inst.isrc_amps = fields[NR_FIELD_ISRC_AMPS] + (size_t)b * nl->n_isrc;
            if (fields[NR_FIELD_VSRC_VOLTS]) inst.vsrc_volts = fields[NR_FIELD_VSRC_VOLTS] + (size_t)b * nl->n_vsrc;
            if (fields[NR_FIELD_CAP_FARADS]) inst.cap_farads = fields[NR_FIELD_CAP_FARADS] + (size_t)b * nl->n_cap;
            if (fields[NR_FIELD_TEMP_K]) inst.ambient_temp_K = fields[NR_FIELD_TEMP_K][b];
        }
        bad |= orc_solve_ac(&inst, t, v_op + (size_t)b * n, o, n_probe, probe,
                            out + (size_t)b * o->n_points * n_probe * 2, &status[b]) != NR_OK;
    }
    return bad ? NR_ERR_INVALID : NR_OK;
}

/* ---- noise analysis (ACSolver::solveNoise, solvers/ac_solver.h:524-583) ----
 * Y(jw) as the AC sweep builds it without an excitation (the reference's
 * noise path calls buildAdmittanceMatrix alone, :559-561).  Noise current
 * sources in computeNoiseSources order (:400-504): every resistor, 4kT/R
 * between its terminals; every diode, 2q|I(vd)| of the clamped diode current
 * at the operating point, anode to cathode; every MOSFET, the channel's
 * 4kT (2/3) gm drain to source when gm > 0 (physics/noise_models.h).  The
 * MOSFET's gm is that of the device the DC solve used, as in the AC sweep.
 * Flicker noise needs KF and C_ox, which default to 0 (netlist/circuit.h:500-
 * 502) and have no netlist field here, so it never applies; BJTs are not part
 * of the netlist.
 *   S_out(f) = sum over sources of |H_k|^2 S_k, added in source order, with
 *   |H|^2 = re re + im im.
 * mode ORC_NOISE_PER_SOURCE restates the reference: one solve of Y V =
 * e_pos - e_neg per source, H_k = V[out] (:565-581).  mode ORC_NOISE_ADJOINT
 * is the GPU's formulation: one solve of Y^T z = e_out per frequency, H_k =
 * z[pos] - z[neg] (the same number up to rounding, since V[out] = e_out^T
 * Y^-1 (e_pos - e_neg)).  The integral is the trapezoid over the frequency
 * points (:585-592). */
static const double NOISE_KB = 1.380649e-23, NOISE_QE = 1.602176634e-19;

typedef struct { int32_t pos, neg; double psd; } noise_src;

static int32_t noise_sources(const nr_netlist* nl, const orc_soa* s, const double* v_op, double temp_K,
                             noise_src* out) {
    const int32_t n = nl->num_nodes;
    int32_t k = 0;
    for (int32_t i = 0; i < nl->n_res; ++i) {
        const double R = nl->res_ohms[i];
        const double psd = R <= 0.0 ? 0.0 : 4.0 * NOISE_KB * temp_K / R;
        if (psd > 0.0) { out[k].pos = nl->res_n1[i]; out[k].neg = nl->res_n2[i]; out[k].psd = psd; ++k; }
    }
    for (int32_t i = 0; i < nl->n_diode; ++i) {
        const int32_t a = nl->dio_anode[i], c = nl->dio_cathode[i];
        const double I = diode_i_clamped(volt(v_op, n, a) - volt(v_op, n, c), nl->dio_is[i], nl->dio_n[i],
                                         nl->dio_vt[i]);
        const double psd = 2.0 * NOISE_QE * fabs(I);
        if (psd > 0.0) { out[k].pos = a; out[k].neg = c; out[k].psd = psd; ++k; }
    }
    int32_t* soa_of = (int32_t*)xmalloc(4 * (size_t)(s->nM + 1));
    for (int32_t p = 0; p < s->nM; ++p) soa_of[s->m_src[p]] = p;
    for (int32_t i = 0; i < nl->n_mos; ++i) {
        const int32_t p = soa_of[i];
        double vgs, vds, ids, gm, gds;
        mos_one(volt(v_op, n, s->mg[p]), volt(v_op, n, s->md[p]), volt(v_op, n, s->ms[p]), s->mW[p], s->mL[p],
                s->mKp[p], s->mVth[p], s->mLam[p], s->mP[p], &vgs, &vds, &ids, &gm, &gds);
        const double psd = gm <= 0.0 ? 0.0 : 4.0 * NOISE_KB * temp_K * (2.0 / 3.0) * gm;
        if (psd > 0.0) { out[k].pos = s->md[p]; out[k].neg = s->ms[p]; out[k].psd = psd; ++k; }
    }
    free(soa_of);
    return k;
}

int orc_solve_noise(const nr_netlist* nl, const orc_topo* t, const double* v_op, const orc_noise_opts* o,
                    double* psd_out, double* vrms, int32_t* status) {
    const int32_t n = nl->num_nodes;
    if (!t || !t->transient || o->n_points <= 0 || o->output_node < 0 || o->output_node > n) return NR_ERR_INVALID;
    orc_soa* s = orc_tensorize(nl);
    const int32_t F = o->n_points;
    orc_ac_opts fo = {o->f_start, o->f_stop, F, 0, 1.0};
    double* freqs = (double*)xmalloc(sizeof(double) * (size_t)F);
    orc_ac_freqs(&fo, freqs);
    double* vre = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* vim = (double*)xmalloc(sizeof(double) * (size_t)t->n_stamps);
    double* are = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    double* aim = (double*)xmalloc(sizeof(double) * (size_t)t->nnz_a);
    cplx* A = (cplx*)xmalloc(sizeof(cplx) * (size_t)t->nnz_a);
    cplx* rhs = (cplx*)xmalloc(sizeof(cplx) * (size_t)n);
    cplx* x = (cplx*)xmalloc(sizeof(cplx) * (size_t)n);
    noise_src* src = (noise_src*)xmalloc(sizeof(noise_src) * (size_t)(nl->n_res + nl->n_diode + nl->n_mos + 1));
    const int32_t ns = noise_sources(nl, s, v_op, o->temp_K, src);
    tran_ctx tc = {1.0, ORC_INT_GEAR1, NULL, NULL, NULL, NULL, NULL, 0};
    emitter e;
    memset(&e, 0, sizeof(e));
    e.vals = vre;
    e.ac = 1;
    stamp_all(nl, s, v_op, 1e-12, &e, &tc);
    int rc = e.nv == t->n_stamps ? NR_OK : NR_ERR_TOPOLOGY;
    orc_assemble(t, vre, are);
    *status = NR_STATUS_CONVERGED;
    const int32_t out = o->output_node;
    for (int32_t f = 0; f < F && rc == NR_OK; ++f) {
        const double omega = 2.0 * 3.14159265358979323846 * freqs[f];
        memset(&e, 0, sizeof(e));
        e.vals = vim;
        e.ac = 2;
        e.omega = omega;
        stamp_all(nl, s, v_op, 1e-12, &e, &tc);
        orc_assemble(t, vim, aim);
        for (int32_t k = 0; k < t->nnz_a; ++k) { A[k].re = are[k]; A[k].im = aim[k]; }
        double total = 0.0;
        if (o->mode == ORC_NOISE_ADJOINT) {
            for (int32_t i = 0; i < n; ++i) { rhs[i].re = 0.0; rhs[i].im = 0.0; }
            if (out > 0) { rhs[out - 1].re = 0.0 + 1.0; rhs[out - 1].im = 0.0 + 0.0; }
            if (lu_multifrontal_c(t, A, rhs, x, 1) != NR_OK) *status = NR_STATUS_SINGULAR;
            for (int32_t k = 0; k < ns; ++k) {
                const cplx zp = src[k].pos > 0 ? x[src[k].pos - 1] : (cplx){0.0, 0.0};
                const cplx zn = src[k].neg > 0 ? x[src[k].neg - 1] : (cplx){0.0, 0.0};
                const double hr = zp.re - zn.re, hi = zp.im - zn.im;
                total = total + (hr * hr + hi * hi) * src[k].psd;
            }
        } else {
            for (int32_t k = 0; k < ns; ++k) {
                for (int32_t i = 0; i < n; ++i) { rhs[i].re = 0.0; rhs[i].im = 0.0; }
                if (src[k].pos > 0) { rhs[src[k].pos - 1].re += 1.0; rhs[src[k].pos - 1].im += 0.0; }
                if (src[k].neg > 0) { rhs[src[k].neg - 1].re -= 1.0; rhs[src[k].neg - 1].im -= 0.0; }
                if (lu_multifrontal_c(t, A, rhs, x, 0) != NR_OK) *status = NR_STATUS_SINGULAR;
                const cplx h = out > 0 ? x[out - 1] : (cplx){0.0, 0.0};
                total = total + (h.re * h.re + h.im * h.im) * src[k].psd;
            }
        }
        psd_out[f] = total;
    }
    double integrated = 0.0;
    for (int32_t i = 1; i < F; ++i) {
        const double df = freqs[i] - freqs[i - 1];
        const double avg = 0.5 * (psd_out[i] + psd_out[i - 1]);
        integrated += avg * df;
    }
    *vrms = sqrt(integrated > 0.0 ? integrated : 0.0);
    free(freqs); free(vre); free(vim); free(are); free(aim); free(A); free(rhs); free(x); free(src);
    orc_soa_free(s);
    return rc;
}

int orc_solve_noise_batch(const nr_netlist* nl, const orc_topo* t, int32_t B, const double* const* fields,
                          const double* v_op, const orc_noise_opts* o, int32_t threads, double* psd_out,
                          double* vrms, int32_t* status) {
    if (!t || B <= 0) return NR_ERR_INVALID;
    const int32_t n = nl->num_nodes;
    int bad = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
    for (int32_t b = 0; b < B; ++b) {
        nr_netlist inst = *nl;
        if (fields) {
            if (fields[NR_FIELD_RES_OHMS]) inst.res_ohms = fields[NR_FIELD_RES_OHMS] + (size_t)b * nl->n_res;
            if (fields[NR_FIELD_DIO_IS]) inst.dio_is = fields[NR_FIELD_DIO_IS] + (size_t)b * nl->n_diode;
            if (fields[NR_FIELD_MOS_KP]) inst.mos_kp = fields[NR_FIELD_MOS_KP] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_MOS_VTO]) inst.mos_vto = fields[NR_FIELD_MOS_VTO] + (size_t)b * nl->n_mos;
            if (fields[NR_FIELD_ISRC_AMPS]) inst.isrc_amps = fields[NR_FIELD_ISRC_AMPS] + (size_t)b * nl->n_isrc;
            if (fields[NR_FIELD_VSRC_VOLTS]) inst.vsrc_volts = fields[NR_FIELD_VSRC_VOLTS] + (size_t)b * nl->n_vsrc;
            if (fields[NR_FIELD_CAP_FARADS]) inst.cap_farads = fields[NR_FIELD_CAP_FARADS] + (size_t)b * nl->n_cap;
            if (fields[NR_FIELD_TEMP_K]) inst.ambient_temp_K = fields[NR_FIELD_TEMP_K][b];
        }
        bad |= orc_solve_noise(&inst, t, v_op + (size_t)b * n, o, psd_out + (size_t)b * o->n_points, &vrms[b],
                               &status[b]) != NR_OK;
    }
    return bad ? NR_ERR_INVALID : NR_OK;
}
