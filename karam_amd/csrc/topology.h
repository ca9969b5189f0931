// topology.h — host-side symbolic analysis of one netlist topology.
//
// Produces everything the batched GPU Newton loop needs that depends only on
// the circuit's structure: the reference SoA device order, the DC stamp
// sequence and its CSR pattern / stamp->slot map, the Min-Fill elimination
// ordering, the filled L+U pattern, relaxed supernodes with their
// frontal index maps, and the per-node incidence lists used by assembly, the
// KCL residual and the pn-junction limiter.  Integer outputs that have a
// reference counterpart are bit-exact with it.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "karam_nr.h"

namespace karam {

// Source of one stamp value (stamp emission of stampSoABlock,
// physics/circuitsim.h:2413-2751, plus addDiagonalConditioning :2857-2860).
enum StampKind : int32_t {
    SK_RES = 0,      // +/- 1/R of resistor idx (netlist order)
    SK_VSRC = 1,     // +/- 1e3 (Norton conductance)
    SK_DIODE = 2,    // +/- g_d of SoA diode idx
    SK_MOS_GDS = 3,  // +/- gds of SoA MOSFET idx
    SK_MOS_GM = 4,   // +/- gm of SoA MOSFET idx
    SK_MOS_GMB = 5,  // +/- gmb (identically 0 for the generic kernel)
    SK_GMIN = 6,     // + gmin
    SK_CAP = 7       // +/- C/h of capacitor idx (transient topologies only)
};

// rhs / KCL contribution sources
enum TermKind : int32_t {
    TK_VSRC = 0,     // rhs: +/- 1e3*V ; KCL: +/- (V - (v+ - v-))/1e-3
    TK_ISRC = 1,     // rhs: -/+ I     ; KCL: -/+ I
    TK_DIODE = 2,    // rhs: -/+ ieq (SoA idx) ; KCL: -/+ i_d (netlist idx)
    TK_MOS = 3,      // rhs: -/+ ieq (SoA idx) ; KCL: -/+ ids (SoA idx)
    TK_RES = 4,      // KCL only: -/+ (v1 - v2)/R
    TK_CAP = 5       // rhs: +/- (C/h) v_prev ; KCL: -/+ (C/h)(v - v_prev)  (transient)
};

struct Stamp {
    int32_t row, col;
    int32_t kind, idx;
    int32_t sign;    // +1 / -1
};

// packed (kind | sign | index) entry used by the device-side lists
inline int32_t pack_term(int32_t kind, int32_t sign, int32_t idx) {
    return (idx << 4) | (kind << 1) | (sign < 0 ? 1 : 0);
}

struct SoA {
    // diodes in tensorizeNetlist order (bucketed by model name)
    std::vector<int32_t> d_a, d_c, d_src;
    std::vector<double> d_is, d_n, d_vt;       // d_vt temperature-scaled
    std::vector<double> d_vtraw;               // unscaled V_T (KCL audit)
    std::vector<double> d_pn_vt, d_pn_isf;     // pnjlim V_T and Is factor
    // MOSFETs in tensorizeNetlist order
    std::vector<int32_t> m_d, m_g, m_s, m_b, m_pmos, m_src;
    std::vector<double> m_w, m_l, m_kp, m_vth, m_lam;
};

struct Supernode {
    int32_t first, ncols;       // permuted column range [first, first+ncols)
    int32_t nrows;              // |R|: rows below the supernode
    int32_t parent;             // supernode index or -1
    int32_t level;              // height above the leaves (0 = leaf)
    int32_t row_off;            // offset into row_idx (R rows, permuted ids)
    int64_t u_off;              // c x (c+m) U block, per-instance offset
    int64_t cb_off;             // m x m contribution block, per-instance offset
    int64_t rcb_off;            // m rhs contribution, per-instance offset
    int32_t a_begin, a_end;     // range into front_a_*
    int32_t child_begin, child_end;  // range into children
};

struct Topology {
    int32_t n = 0;
    SoA soa;
    // netlist-order component tables needed on the device
    int32_t n_res = 0, n_vsrc = 0, n_isrc = 0, n_cap = 0, n_dio_nl = 0, n_mos_nl = 0;
    std::vector<int32_t> res_n1, res_n2, vsrc_pos, vsrc_neg, isrc_pos, isrc_neg, dio_anode_nl;
    std::vector<double> res_ohms, vsrc_volts, isrc_amps;
    std::vector<int32_t> cap_n1, cap_n2, vsrc_wave, isrc_wave;   // NR_WAVE_* per source
    std::vector<double> cap_farads, vsrc_wave_par, isrc_wave_par;  // 7 values per source
    // transient topology: capacitor companion stamps are part of the pattern
    bool transient = false;

    // stamps in emission order, CSR pattern of A and the stamp map
    std::vector<Stamp> stamps;
    std::vector<int32_t> row_ptr, col_idx, stamp_slot;
    int32_t nnz_a = 0;
    // per-slot stamp lists (emission order): linear part then nonlinear part
    std::vector<int32_t> lin_ptr, lin_term;     // SK_RES / SK_VSRC / SK_CAP
    std::vector<int32_t> nl_ptr, nl_term;       // SK_DIODE / SK_MOS_* (gmb skipped)
    std::vector<int32_t> slot_is_diag;          // gmin added last
    // per-node rhs lists (emission order)
    std::vector<int32_t> rhs_lin_ptr, rhs_lin_term;   // TK_VSRC / TK_ISRC / TK_CAP
    std::vector<int32_t> rhs_nl_ptr, rhs_nl_term;     // TK_DIODE / TK_MOS (SoA)
    // per-node KCL lists (netlist order of calculatePhysicalResiduals)
    std::vector<int32_t> kcl_ptr, kcl_term;
    // pnjlim: diodes (SoA order) grouped into dependency levels
    std::vector<int32_t> pn_level_ptr, pn_diode;
    // Min-Fill
    std::vector<int32_t> ordering_raw;
    int32_t n_vars = 0;
    int32_t treewidth = 0;
    std::vector<int32_t> perm, iperm;
    // filled pattern (permuted space)
    std::vector<int32_t> frow_ptr, fcol_idx;
    int32_t nnz_f = 0;
    int32_t fill_in = 0;
    // supernodes
    std::vector<Supernode> sn;
    std::vector<int32_t> row_idx;            // concatenated R rows (permuted)
    std::vector<int32_t> children;           // concatenated child lists
    std::vector<int32_t> relmap;             // per supernode: R rows -> parent front index
    std::vector<int32_t> col2sn;
    std::vector<int32_t> level_ptr, level_sn;  // supernodes grouped by level
    int32_t n_levels = 0, max_front = 0;
    int64_t u_size = 0, cb_size = 0, rcb_size = 0;
    int64_t lu_flops = 0;
    // A entries per supernode: CSR slot and front-local (row, col)
    std::vector<int32_t> front_a_slot, front_a_li, front_a_lj;
    double analyse_ms = 0.0;
};

// Build; returns NR_OK or an error code with a message in `err`.  A
// transient topology also carries the capacitor companion stamps.
int build_topology(const nr_netlist* nl, Topology& t, std::string& err, bool transient = false);

// tensorizeNetlist order and parameters (tensors/physics_tensors.cpp:305-493).
void tensorize(const nr_netlist* nl, SoA& s);

}  // namespace karam
