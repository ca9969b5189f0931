// topology.cpp — symbolic analysis of one netlist topology (host side).
//
// Reference stages reproduced here (integer results bit-exact):
//   tensorizeNetlist            tensors/physics_tensors.cpp:305-493
//   stampSoABlock emission      physics/circuitsim.h:2413-2751 (+ :2857-2860)
//   createCsr / buildPattern    math/linalg.h:281-431 (stamp map :441-465)
//   FactorGraph + Min-Fill      infrastructure/factor_graph.cpp:9-162,
//                               infrastructure/treewidth_analyzer.cpp:14-141
//   symbolicFactorize           math/sparse_lu.cpp:14-128
// Min-Fill is computed incrementally (fill counts updated only where an
// elimination can change them) with the reference's selection rule — min
// fill, ties to the smallest id, duplicate ids as coded — so it yields
// the same ordering as the reference's full rescan.  The fill pattern comes
// from the elimination tree (column structure = own lower entries united with
// the children's structures), which equals the reference's simulated graph
// elimination.
#include "topology.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>

namespace karam {

namespace {

const char* model_name(const nr_netlist* nl, int32_t idx, const char* dflt) {
    if (idx < 0 || idx >= nl->n_models || !nl->models || !nl->models[idx].name) return dflt;
    return nl->models[idx].name;
}

const nr_model_card* find_card(const nr_netlist* nl, const std::string& name) {
    for (int32_t i = 0; i < nl->n_models; ++i) {
        const nr_model_card& c = nl->models[i];
        if (c.type && c.name && name == c.name) return &c;
    }
    return nullptr;
}

}  // namespace

void tensorize(const nr_netlist* nl, SoA& s) {
    s = SoA();
    const double tscale = nl->ambient_temp_K / 300.15;
    // Diodes, bucketed by model name (std::map order), netlist order inside.
    std::vector<std::pair<std::string, int32_t>> items;
    for (int32_t i = 0; i < nl->n_diode; ++i)
        items.emplace_back(model_name(nl, nl->dio_model ? nl->dio_model[i] : -1, "diode_generic"), i);
    std::stable_sort(items.begin(), items.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& it : items) {
        int32_t i = it.second;
        s.d_src.push_back(i);
        s.d_a.push_back(nl->dio_anode[i]);
        s.d_c.push_back(nl->dio_cathode[i]);
        s.d_is.push_back(nl->dio_is[i]);
        s.d_n.push_back(nl->dio_n[i]);
        s.d_vt.push_back(nl->dio_vt[i] * tscale);
        s.d_vtraw.push_back(nl->dio_vt[i]);
        // applyPnJunctionLimits scales V_T and Is once more by the ambient
        // temperature (solvers/circuitsim_solver.cpp:127-129)
        double vt = s.d_vt.back() * tscale;
        s.d_pn_vt.push_back(vt);
        s.d_pn_isf.push_back(std::pow(tscale, 3.0) * std::exp(1.11 / (vt / tscale) - 1.11 / vt));
    }
    // MOSFETs: bucket by model name; card parameters or tensorizer defaults.
    items.clear();
    for (int32_t i = 0; i < nl->n_mos; ++i)
        items.emplace_back(model_name(nl, nl->mos_model ? nl->mos_model[i] : -1, ""), i);
    std::stable_sort(items.begin(), items.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    bool any_card = false;
    for (int32_t i = 0; i < nl->n_models; ++i) any_card |= (nl->models[i].type != nullptr);
    for (auto& it : items) {
        int32_t i = it.second;
        const std::string& name = it.first;
        bool pm = false;
        double kp = 200e-6, vth = 0.7, lam = 0.02;
        if (any_card) {
            const nr_model_card* c = find_card(nl, name);
            if (c) {
                pm = std::strcmp(c->type, "PMOS") == 0;
                kp = c->has_kp ? c->kp : (pm ? 100e-6 : 200e-6);
                vth = c->has_vto ? c->vto : (pm ? -0.7 : 0.7);
                lam = c->has_lambda ? c->lambda : 0.02;
            } else {
                pm = name.find("PMOS") != std::string::npos || name.find("pmos") != std::string::npos;
                if (pm) { kp = 100e-6; vth = -0.7; }
            }
        }
        if (nl->mos_kp) kp = nl->mos_kp[i];     // per-device MC override
        if (nl->mos_vto) vth = nl->mos_vto[i];
        s.m_src.push_back(i);
        s.m_d.push_back(nl->mos_d[i]); s.m_g.push_back(nl->mos_g[i]);
        s.m_s.push_back(nl->mos_s[i]); s.m_b.push_back(nl->mos_b[i]);
        s.m_w.push_back(nl->mos_w[i]); s.m_l.push_back(nl->mos_l[i]);
        s.m_kp.push_back(kp); s.m_vth.push_back(vth); s.m_lam.push_back(lam);
        s.m_pmos.push_back(pm ? 1 : 0);
    }
}

namespace {

// DC stamp emission order of stampSoABlock (h = 0) + gmin diagonal.
void emit_stamps(const nr_netlist* nl, const SoA& s, std::vector<Stamp>& out, bool transient) {
    out.clear();
    auto st = [&](int32_t r, int32_t c, int32_t k, int32_t idx, int32_t sg) {
        out.push_back(Stamp{r, c, k, idx, sg});
    };
    for (int32_t i = 0; i < nl->n_res; ++i) {
        int32_t a = nl->res_n1[i], b = nl->res_n2[i];
        if (a != 0 && b != 0) {
            st(a - 1, a - 1, SK_RES, i, 1); st(b - 1, b - 1, SK_RES, i, 1);
            st(a - 1, b - 1, SK_RES, i, -1); st(b - 1, a - 1, SK_RES, i, -1);
        } else if (a != 0) st(a - 1, a - 1, SK_RES, i, 1);
        else if (b != 0) st(b - 1, b - 1, SK_RES, i, 1);
    }
    for (int32_t i = 0; i < nl->n_vsrc; ++i) {
        int32_t p = nl->vsrc_pos[i], q = nl->vsrc_neg[i];
        if (p != 0 && q != 0) {
            st(p - 1, p - 1, SK_VSRC, i, 1); st(q - 1, q - 1, SK_VSRC, i, 1);
            st(p - 1, q - 1, SK_VSRC, i, -1); st(q - 1, p - 1, SK_VSRC, i, -1);
        } else if (p != 0) st(p - 1, p - 1, SK_VSRC, i, 1);
        else if (q != 0) st(q - 1, q - 1, SK_VSRC, i, 1);
    }
    // capacitor companions follow the sources (physics/circuitsim.h:2479-2505)
    for (int32_t i = 0; transient && i < nl->n_cap; ++i) {
        int32_t a = nl->cap_n1[i], b = nl->cap_n2[i];
        if (a != 0 && b != 0) {
            st(a - 1, a - 1, SK_CAP, i, 1); st(b - 1, b - 1, SK_CAP, i, 1);
            st(a - 1, b - 1, SK_CAP, i, -1); st(b - 1, a - 1, SK_CAP, i, -1);
        } else if (a != 0) st(a - 1, a - 1, SK_CAP, i, 1);
        else if (b != 0) st(b - 1, b - 1, SK_CAP, i, 1);
    }
    const int32_t nD = (int32_t)s.d_a.size();
    for (int32_t i = 0; i < nD; ++i) {
        int32_t a = s.d_a[i], c = s.d_c[i];
        if (a != 0 && c != 0) {
            st(a - 1, a - 1, SK_DIODE, i, 1); st(c - 1, c - 1, SK_DIODE, i, 1);
            st(a - 1, c - 1, SK_DIODE, i, -1); st(c - 1, a - 1, SK_DIODE, i, -1);
        } else if (a != 0) st(a - 1, a - 1, SK_DIODE, i, 1);
        else if (c != 0) st(c - 1, c - 1, SK_DIODE, i, 1);
    }
    const int32_t nM = (int32_t)s.m_d.size();
    for (int32_t i = 0; i < nM; ++i) {
        int32_t d = s.m_d[i], g = s.m_g[i], so = s.m_s[i], b = s.m_b[i];
        if (d != 0 && so != 0) {
            st(d - 1, d - 1, SK_MOS_GDS, i, 1); st(so - 1, so - 1, SK_MOS_GDS, i, 1);
            st(d - 1, so - 1, SK_MOS_GDS, i, -1); st(so - 1, d - 1, SK_MOS_GDS, i, -1);
        } else if (d != 0) st(d - 1, d - 1, SK_MOS_GDS, i, 1);
        else if (so != 0) st(so - 1, so - 1, SK_MOS_GDS, i, 1);
        if (d != 0) {
            if (g != 0) st(d - 1, g - 1, SK_MOS_GM, i, 1);
            if (so != 0) st(d - 1, so - 1, SK_MOS_GM, i, -1);
        }
        if (so != 0) {
            if (g != 0) st(so - 1, g - 1, SK_MOS_GM, i, -1);
            st(so - 1, so - 1, SK_MOS_GM, i, 1);
        }
        if (d != 0) {
            if (b != 0) st(d - 1, b - 1, SK_MOS_GMB, i, 1);
            if (so != 0) st(d - 1, so - 1, SK_MOS_GMB, i, -1);
        }
        if (so != 0) {
            if (b != 0) st(so - 1, b - 1, SK_MOS_GMB, i, -1);
            st(so - 1, so - 1, SK_MOS_GMB, i, 1);
        }
    }
    for (int32_t i = 0; i < nl->num_nodes; ++i) st(i, i, SK_GMIN, 0, 1);
}

// Incremental exact Min-Fill over the factor graph.
struct MinFill {
    std::vector<std::vector<int32_t>> adj;   // sorted neighbour lists
    std::vector<uint32_t> fill;
    std::vector<int32_t> mult;
    std::vector<char> elim;
    std::vector<int32_t> mark;
    int32_t token = 0;

    bool has(int32_t u, int32_t w) const {
        const auto& a = adj[u];
        return std::binary_search(a.begin(), a.end(), w);
    }
    void add_edge(int32_t u, int32_t w) {
        auto& a = adj[u];
        a.insert(std::lower_bound(a.begin(), a.end(), w), w);
        auto& b = adj[w];
        b.insert(std::lower_bound(b.begin(), b.end(), u), u);
    }
    uint32_t compute_fill(int32_t u) {
        const auto& nu = adj[u];
        int64_t d = (int64_t)nu.size();
        if (d < 2) return 0;
        ++token;
        for (int32_t w : nu) mark[w] = token;
        int64_t edges = 0;
        for (int32_t w : nu)
            for (int32_t y : adj[w])
                if (y > w && mark[y] == token) ++edges;
        return (uint32_t)(d * (d - 1) / 2 - edges);
    }
};

}  // namespace

int build_topology(const nr_netlist* nl, Topology& t, std::string& err, bool transient) {
    auto t0 = std::chrono::steady_clock::now();
    if (!nl || nl->num_nodes <= 0) { err = "netlist has no nodes"; return NR_ERR_INVALID; }
    const int32_t n = nl->num_nodes;
    auto bad_node = [&](int32_t x) { return x < 0 || x > n; };
    for (int32_t i = 0; i < nl->n_res; ++i)
        if (bad_node(nl->res_n1[i]) || bad_node(nl->res_n2[i]) || !(nl->res_ohms[i] != 0.0)) {
            err = "resistor node out of range or zero resistance"; return NR_ERR_INVALID;
        }
    for (int32_t i = 0; i < nl->n_vsrc; ++i)
        if (bad_node(nl->vsrc_pos[i]) || bad_node(nl->vsrc_neg[i])) { err = "vsource node out of range"; return NR_ERR_INVALID; }
    for (int32_t i = 0; i < nl->n_isrc; ++i)
        if (bad_node(nl->isrc_pos[i]) || bad_node(nl->isrc_neg[i])) { err = "isource node out of range"; return NR_ERR_INVALID; }
    auto bad_wave = [](int32_t count, const int32_t* kind) {
        for (int32_t i = 0; kind && i < count; ++i)
            if (kind[i] < NR_WAVE_DC || kind[i] > NR_WAVE_SINE) return true;
        return false;
    };
    if (bad_wave(nl->n_vsrc, nl->vsrc_wave) || bad_wave(nl->n_isrc, nl->isrc_wave)) {
        err = "unknown source waveform kind"; return NR_ERR_INVALID;
    }
    for (int32_t i = 0; i < nl->n_cap; ++i)
        if (bad_node(nl->cap_n1[i]) || bad_node(nl->cap_n2[i])) { err = "capacitor node out of range"; return NR_ERR_INVALID; }
    for (int32_t i = 0; i < nl->n_diode; ++i)
        if (bad_node(nl->dio_anode[i]) || bad_node(nl->dio_cathode[i])) { err = "diode node out of range"; return NR_ERR_INVALID; }
    for (int32_t i = 0; i < nl->n_mos; ++i)
        if (bad_node(nl->mos_d[i]) || bad_node(nl->mos_g[i]) || bad_node(nl->mos_s[i]) || bad_node(nl->mos_b[i])) {
            err = "mosfet node out of range"; return NR_ERR_INVALID;
        }
    t = Topology();
    t.n = n;
    t.n_res = nl->n_res; t.n_vsrc = nl->n_vsrc; t.n_isrc = nl->n_isrc; t.n_cap = nl->n_cap;
    t.n_dio_nl = nl->n_diode; t.n_mos_nl = nl->n_mos;
    t.res_n1.assign(nl->res_n1, nl->res_n1 + nl->n_res);
    t.res_n2.assign(nl->res_n2, nl->res_n2 + nl->n_res);
    t.res_ohms.assign(nl->res_ohms, nl->res_ohms + nl->n_res);
    t.vsrc_pos.assign(nl->vsrc_pos, nl->vsrc_pos + nl->n_vsrc);
    t.vsrc_neg.assign(nl->vsrc_neg, nl->vsrc_neg + nl->n_vsrc);
    t.vsrc_volts.assign(nl->vsrc_volts, nl->vsrc_volts + nl->n_vsrc);
    t.isrc_pos.assign(nl->isrc_pos, nl->isrc_pos + nl->n_isrc);
    t.isrc_neg.assign(nl->isrc_neg, nl->isrc_neg + nl->n_isrc);
    t.isrc_amps.assign(nl->isrc_amps, nl->isrc_amps + nl->n_isrc);
    t.dio_anode_nl.assign(nl->dio_anode, nl->dio_anode + nl->n_diode);
    t.cap_n1.assign(nl->cap_n1, nl->cap_n1 + nl->n_cap);
    t.cap_n2.assign(nl->cap_n2, nl->cap_n2 + nl->n_cap);
    t.cap_farads.assign(nl->cap_farads, nl->cap_farads + nl->n_cap);
    auto waves = [](int32_t count, const int32_t* kind, const double* par, std::vector<int32_t>& k,
                    std::vector<double>& p) {
        k.assign(count, NR_WAVE_DC);
        p.assign((size_t)count * 7, 0.0);
        for (int32_t i = 0; i < count && kind && par; ++i) {
            k[i] = kind[i];
            for (int q = 0; q < 7; ++q) p[(size_t)i * 7 + q] = par[(size_t)i * 7 + q];
        }
    };
    waves(nl->n_vsrc, nl->vsrc_wave, nl->vsrc_wave_par, t.vsrc_wave, t.vsrc_wave_par);
    waves(nl->n_isrc, nl->isrc_wave, nl->isrc_wave_par, t.isrc_wave, t.isrc_wave_par);
    t.transient = transient;
    tensorize(nl, t.soa);
    const SoA& s = t.soa;
    const int32_t nD = (int32_t)s.d_a.size(), nM = (int32_t)s.m_d.size();

    // ---- stamps, CSR pattern, stamp map -------------------------------------
    emit_stamps(nl, s, t.stamps, transient);
    const int32_t S = (int32_t)t.stamps.size();
    std::vector<int32_t> order(S);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const Stamp &x = t.stamps[a], &y = t.stamps[b];
        return x.row < y.row || (x.row == y.row && x.col < y.col);
    });
    t.stamp_slot.assign(S, -1);
    t.row_ptr.assign(n + 1, 0);
    t.col_idx.clear();
    int64_t last = -1;
    for (int32_t k = 0; k < S; ++k) {
        const Stamp& x = t.stamps[order[k]];
        int64_t key = ((int64_t)x.row << 32) | (uint32_t)x.col;
        if (key != last) {
            t.col_idx.push_back(x.col);
            t.row_ptr[x.row + 1]++;
            last = key;
        }
        t.stamp_slot[order[k]] = (int32_t)t.col_idx.size() - 1;
    }
    for (int32_t i = 0; i < n; ++i) t.row_ptr[i + 1] += t.row_ptr[i];
    t.nnz_a = (int32_t)t.col_idx.size();

    // per-slot stamp lists in emission order (order[] is stable, so within a
    // slot the stamps are already in emission order)
    t.lin_ptr.assign(t.nnz_a + 1, 0);
    t.nl_ptr.assign(t.nnz_a + 1, 0);
    t.slot_is_diag.assign(t.nnz_a, 0);
    for (int32_t k = 0; k < S; ++k) {
        const Stamp& x = t.stamps[k];
        int32_t sl = t.stamp_slot[k];
        if (x.kind == SK_RES || x.kind == SK_VSRC || x.kind == SK_CAP) t.lin_ptr[sl + 1]++;
        else if (x.kind == SK_DIODE || x.kind == SK_MOS_GDS || x.kind == SK_MOS_GM) t.nl_ptr[sl + 1]++;
        else if (x.kind == SK_GMIN) t.slot_is_diag[sl] = 1;
    }
    for (int32_t i = 0; i < t.nnz_a; ++i) { t.lin_ptr[i + 1] += t.lin_ptr[i]; t.nl_ptr[i + 1] += t.nl_ptr[i]; }
    t.lin_term.assign(t.lin_ptr[t.nnz_a], 0);
    t.nl_term.assign(t.nl_ptr[t.nnz_a], 0);
    {
        std::vector<int32_t> fl(t.lin_ptr.begin(), t.lin_ptr.end() - 1), fn(t.nl_ptr.begin(), t.nl_ptr.end() - 1);
        for (int32_t k = 0; k < S; ++k) {      // emission order
            const Stamp& x = t.stamps[k];
            int32_t sl = t.stamp_slot[k];
            if (x.kind == SK_RES || x.kind == SK_VSRC || x.kind == SK_CAP)
                t.lin_term[fl[sl]++] = pack_term(x.kind, x.sign, x.idx);
            else if (x.kind == SK_DIODE || x.kind == SK_MOS_GDS || x.kind == SK_MOS_GM)
                t.nl_term[fn[sl]++] = pack_term(x.kind, x.sign, x.idx);
        }
    }

    // ---- per-node rhs lists (stampSoABlock rhs order) ---------------------
    {
        std::vector<std::vector<int32_t>> lin(n), nlt(n);
        for (int32_t i = 0; i < nl->n_vsrc; ++i) {
            int32_t p = nl->vsrc_pos[i], q = nl->vsrc_neg[i];
            if (p != 0) lin[p - 1].push_back(pack_term(TK_VSRC, 1, i));
            if (q != 0) lin[q - 1].push_back(pack_term(TK_VSRC, -1, i));
        }
        for (int32_t i = 0; i < nl->n_isrc; ++i) {
            int32_t p = nl->isrc_pos[i], q = nl->isrc_neg[i];
            if (p != 0) lin[p - 1].push_back(pack_term(TK_ISRC, -1, i));
            if (q != 0) lin[q - 1].push_back(pack_term(TK_ISRC, 1, i));
        }
        for (int32_t i = 0; transient && i < nl->n_cap; ++i) {   // history current
            int32_t a = nl->cap_n1[i], b = nl->cap_n2[i];
            if (a != 0) lin[a - 1].push_back(pack_term(TK_CAP, 1, i));
            if (b != 0) lin[b - 1].push_back(pack_term(TK_CAP, -1, i));
        }
        for (int32_t i = 0; i < nD; ++i) {
            if (s.d_a[i] != 0) nlt[s.d_a[i] - 1].push_back(pack_term(TK_DIODE, -1, i));
            if (s.d_c[i] != 0) nlt[s.d_c[i] - 1].push_back(pack_term(TK_DIODE, 1, i));
        }
        for (int32_t i = 0; i < nM; ++i) {
            if (s.m_d[i] != 0) nlt[s.m_d[i] - 1].push_back(pack_term(TK_MOS, -1, i));
            if (s.m_s[i] != 0) nlt[s.m_s[i] - 1].push_back(pack_term(TK_MOS, 1, i));
        }
        t.rhs_lin_ptr.assign(n + 1, 0); t.rhs_nl_ptr.assign(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) {
            t.rhs_lin_ptr[i + 1] = t.rhs_lin_ptr[i] + (int32_t)lin[i].size();
            t.rhs_nl_ptr[i + 1] = t.rhs_nl_ptr[i] + (int32_t)nlt[i].size();
            t.rhs_lin_term.insert(t.rhs_lin_term.end(), lin[i].begin(), lin[i].end());
            t.rhs_nl_term.insert(t.rhs_nl_term.end(), nlt[i].begin(), nlt[i].end());
        }
    }

    // ---- per-node KCL lists (calculatePhysicalResiduals order,
    //      physics/circuitsim.h:1250-1671, netlist order; device terms use
    //      SoA indices) ----------------------------------------------------
    {
        std::vector<int32_t> dpos(nD), mpos(nM);
        for (int32_t p = 0; p < nD; ++p) dpos[s.d_src[p]] = p;
        for (int32_t p = 0; p < nM; ++p) mpos[s.m_src[p]] = p;
        std::vector<std::vector<int32_t>> k(n);
        auto add = [&](int32_t node, int32_t term) { if (node > 0 && node <= n) k[node - 1].push_back(term); };
        for (int32_t i = 0; i < nl->n_res; ++i) {
            add(nl->res_n1[i], pack_term(TK_RES, -1, i)); add(nl->res_n2[i], pack_term(TK_RES, 1, i));
        }
        for (int32_t i = 0; i < nl->n_vsrc; ++i) {
            add(nl->vsrc_pos[i], pack_term(TK_VSRC, 1, i)); add(nl->vsrc_neg[i], pack_term(TK_VSRC, -1, i));
        }
        for (int32_t i = 0; i < nl->n_isrc; ++i) {
            add(nl->isrc_pos[i], pack_term(TK_ISRC, -1, i)); add(nl->isrc_neg[i], pack_term(TK_ISRC, 1, i));
        }
        for (int32_t i = 0; i < nl->n_diode; ++i) {
            add(nl->dio_anode[i], pack_term(TK_DIODE, -1, dpos[i])); add(nl->dio_cathode[i], pack_term(TK_DIODE, 1, dpos[i]));
        }
        for (int32_t i = 0; i < nl->n_mos; ++i) {
            add(nl->mos_d[i], pack_term(TK_MOS, -1, mpos[i])); add(nl->mos_s[i], pack_term(TK_MOS, 1, mpos[i]));
        }
        for (int32_t i = 0; transient && i < nl->n_cap; ++i) {   // displacement current (:1460-1490)
            add(nl->cap_n1[i], pack_term(TK_CAP, -1, i)); add(nl->cap_n2[i], pack_term(TK_CAP, 1, i));
        }
        t.kcl_ptr.assign(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) {
            t.kcl_ptr[i + 1] = t.kcl_ptr[i] + (int32_t)k[i].size();
            t.kcl_term.insert(t.kcl_term.end(), k[i].begin(), k[i].end());
        }
    }

    // ---- pnjlim dependency levels (sequential SoA semantics of
    //      applyPnJunctionLimits, solvers/circuitsim_solver.cpp:118-169) ----
    {
        std::vector<int32_t> last_lv(n + 1, -1), lv(nD, -1);
        int32_t nlev = 0;
        for (int32_t i = 0; i < nD; ++i) {
            int32_t p = s.d_a[i], q = s.d_c[i];
            if ((p <= 0 && q <= 0) || p > n || q > n) continue;   // skipped by the reference
            int32_t l = 0;
            if (p > 0) l = std::max(l, last_lv[p] + 1);
            if (q > 0) l = std::max(l, last_lv[q] + 1);
            lv[i] = l;
            if (p > 0) last_lv[p] = l;
            if (q > 0) last_lv[q] = l;
            nlev = std::max(nlev, l + 1);
        }
        t.pn_level_ptr.assign(nlev + 1, 0);
        for (int32_t i = 0; i < nD; ++i) if (lv[i] >= 0) t.pn_level_ptr[lv[i] + 1]++;
        for (int32_t l = 0; l < nlev; ++l) t.pn_level_ptr[l + 1] += t.pn_level_ptr[l];
        t.pn_diode.assign(t.pn_level_ptr[nlev], 0);
        std::vector<int32_t> f(t.pn_level_ptr.begin(), t.pn_level_ptr.end());
        for (int32_t i = 0; i < nD; ++i) if (lv[i] >= 0) t.pn_diode[f[lv[i]]++] = i;
    }

    // ---- factor graph + Min-Fill -------------------------------------------
    {
        const int32_t nbranch = nl->n_vsrc;     // + inductors (none in this ABI)
        const int32_t maxid = n + nbranch;
        MinFill mf;
        mf.adj.assign(maxid + 1, {});
        mf.mult.assign(maxid + 1, 0);
        mf.elim.assign(maxid + 1, 0);
        mf.mark.assign(maxid + 1, 0);
        std::vector<char> touched(maxid + 1, 0);
        std::vector<std::pair<int32_t, int32_t>> edges;
        auto factor = [&](std::initializer_list<int32_t> ids, int32_t extra) {
            int32_t fv[8]; int32_t k = 0;
            for (int32_t x : ids) if (x > 0) { fv[k++] = x; touched[x] = 1; }
            if (extra >= 0) fv[k++] = extra;
            for (int32_t i = 0; i < k; ++i)
                for (int32_t j = i + 1; j < k; ++j)
                    if (fv[i] != fv[j]) { edges.emplace_back(fv[i], fv[j]); edges.emplace_back(fv[j], fv[i]); }
        };
        for (int32_t i = 0; i < nD; ++i) factor({s.d_a[i], s.d_c[i]}, -1);
        for (int32_t i = 0; i < nM; ++i) factor({s.m_d[i], s.m_g[i], s.m_s[i], s.m_b[i]}, -1);
        for (int32_t i = 0; i < nl->n_res; ++i) factor({nl->res_n1[i], nl->res_n2[i]}, -1);
        for (int32_t i = 0; i < nl->n_cap; ++i) factor({nl->cap_n1[i], nl->cap_n2[i]}, -1);
        for (int32_t i = 0; i < nl->n_vsrc; ++i) factor({nl->vsrc_pos[i], nl->vsrc_neg[i]}, n + i);
        for (int32_t i = 0; i < nl->n_isrc; ++i) factor({nl->isrc_pos[i], nl->isrc_neg[i]}, -1);
        std::sort(edges.begin(), edges.end());
        edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
        for (auto& e : edges) mf.adj[e.first].push_back(e.second);   // sorted per vertex
        for (int32_t id = 1; id <= n; ++id) if (touched[id]) mf.mult[id]++;
        for (int32_t k = 0; k < nbranch; ++k) mf.mult[n + k]++;
        int32_t nvars = 0;
        for (int32_t id = 0; id <= maxid; ++id) nvars += mf.mult[id];
        t.n_vars = nvars;
        mf.fill.assign(maxid + 1, 0);
        std::set<std::pair<uint32_t, int32_t>> pq;
        for (int32_t id = 0; id <= maxid; ++id)
            if (mf.mult[id] > 0) { mf.fill[id] = mf.compute_fill(id); pq.insert({mf.fill[id], id}); }
        uint32_t maxClique = 0;
        t.ordering_raw.reserve(nvars);
        std::vector<int32_t> N;
        std::vector<std::pair<int32_t, int32_t>> newe;
        std::vector<char> inN(maxid + 1, 0);
        while (!pq.empty()) {
            auto top = *pq.begin();
            pq.erase(pq.begin());
            const int32_t v = top.second;
            t.ordering_raw.push_back(v);
            if (mf.elim[v]) {                     // duplicate id after its elimination
                maxClique = std::max<uint32_t>(maxClique, 1);
                if (--mf.mult[v] > 0) pq.insert({0u, v});
                continue;
            }
            N = mf.adj[v];
            maxClique = std::max<uint32_t>(maxClique, (uint32_t)N.size() + 1);
            newe.clear();
            for (size_t i = 0; i < N.size(); ++i)
                for (size_t j = i + 1; j < N.size(); ++j)
                    if (!mf.has(N[i], N[j])) { mf.add_edge(N[i], N[j]); newe.emplace_back(N[i], N[j]); }
            for (int32_t u : N) {
                auto& a = mf.adj[u];
                a.erase(std::lower_bound(a.begin(), a.end(), v));
            }
            mf.adj[v].clear();
            mf.elim[v] = 1;
            if (--mf.mult[v] > 0) pq.insert({0u, v});
            for (int32_t u : N) inN[u] = 1;
            // vertices outside N(v) keep their neighbourhood; every new edge
            // inside it makes one of their neighbour pairs adjacent
            for (auto& e : newe) {
                const auto& A = mf.adj[e.first];
                const auto& Bv = mf.adj[e.second];
                size_t i = 0, j = 0;
                while (i < A.size() && j < Bv.size()) {
                    if (A[i] < Bv[j]) ++i;
                    else if (Bv[j] < A[i]) ++j;
                    else {
                        int32_t x = A[i];
                        if (!inN[x] && !mf.elim[x] && mf.mult[x] > 0) {
                            pq.erase({mf.fill[x], x});
                            mf.fill[x] -= 1;
                            pq.insert({mf.fill[x], x});
                        }
                        ++i; ++j;
                    }
                }
            }
            for (int32_t u : N) {
                inN[u] = 0;
                if (mf.elim[u] || mf.mult[u] <= 0) continue;
                pq.erase({mf.fill[u], u});
                mf.fill[u] = mf.compute_fill(u);
                pq.insert({mf.fill[u], u});
            }
        }
        t.treewidth = maxClique > 0 ? (int32_t)maxClique - 1 : 0;
    }

    // ---- projection of the ordering onto matrix rows ----------------------------
    {
        std::vector<char> seen(n + 1, 0);
        t.perm.clear();
        for (int32_t id : t.ordering_raw)
            if (id >= 1 && id <= n && !seen[id]) { seen[id] = 1; t.perm.push_back(id - 1); }
        for (int32_t id = 1; id <= n; ++id) if (!seen[id]) t.perm.push_back(id - 1);
        t.iperm.assign(n, 0);
        for (int32_t i = 0; i < n; ++i) t.iperm[t.perm[i]] = i;
    }

    // ---- symbolic factorisation via the elimination tree ----------------------
    std::vector<std::vector<int32_t>> cs(n);    // struct(L(:,j)) \ {j}, sorted
    std::vector<int32_t> parent(n, -1);
    {
        std::vector<std::vector<int32_t>> lower(n);   // A entries below the diagonal per column
        int64_t orig_edges = 0;
        std::vector<std::pair<int32_t, int32_t>> pe;
        for (int32_t r = 0; r < n; ++r)
            for (int32_t k = t.row_ptr[r]; k < t.row_ptr[r + 1]; ++k) {
                int32_t c = t.col_idx[k];
                if (c == r) continue;
                int32_t pr = t.iperm[r], pc = t.iperm[c];
                pe.emplace_back(std::min(pr, pc), std::max(pr, pc));
            }
        std::sort(pe.begin(), pe.end());
        pe.erase(std::unique(pe.begin(), pe.end()), pe.end());
        orig_edges = (int64_t)pe.size();
        for (auto& e : pe) lower[e.first].push_back(e.second);
        std::vector<std::vector<int32_t>> kids(n);
        int64_t filled_edges = 0;
        std::vector<int32_t> tmp;
        for (int32_t j = 0; j < n; ++j) {
            std::vector<int32_t>& st = cs[j];
            st = lower[j];
            for (int32_t c : kids[j]) {
                tmp.clear();
                std::set_union(st.begin(), st.end(), cs[c].begin() + 1, cs[c].end(), std::back_inserter(tmp));
                st.swap(tmp);
            }
            if (!st.empty()) { parent[j] = st[0]; kids[st[0]].push_back(j); }
            filled_edges += (int64_t)st.size();
        }
        t.fill_in = (int32_t)(filled_edges - orig_edges);
        // filled CSR rows: {j < i : i in cs[j]} + {i} + cs[i]
        std::vector<std::vector<int32_t>> low(n);
        for (int32_t j = 0; j < n; ++j) for (int32_t i : cs[j]) low[i].push_back(j);
        t.frow_ptr.assign(n + 1, 0);
        t.fcol_idx.clear();
        for (int32_t i = 0; i < n; ++i) {
            t.fcol_idx.insert(t.fcol_idx.end(), low[i].begin(), low[i].end());
            t.fcol_idx.push_back(i);
            t.fcol_idx.insert(t.fcol_idx.end(), cs[i].begin(), cs[i].end());
            t.frow_ptr[i + 1] = (int32_t)t.fcol_idx.size();
        }
        t.nnz_f = (int32_t)t.fcol_idx.size();
    }

    // ---- relaxed supernodes -------------------------------------------------------
    // Consecutive columns along an elimination-tree chain share a front.
    // Scanning from the last column down, column j-1 joins the supernode
    // [j, last] when j is its parent and either the pair is fundamental
    // (|struct(j-1)| = |struct(j)| + 1, j has no other child) or the merged
    // front's explicit zeros stay within a fifth of its L entries, at front
    // order <= AMALG_MAX_ORDER and at most AMALG_MAX_PIVOTS pivots (the
    // elimination of a front's pivot block is sequential in its pivots).  A column's structure lies within the later
    // columns of its supernode and the last column's rows (etree property),
    // so the front [first..last] + struct(last) holds every entry.
    {
        constexpr int64_t AMALG_MAX_ORDER = 512, AMALG_MAX_PIVOTS = 16;
        std::vector<int32_t> nchild(n, 0);
        for (int32_t j = 0; j < n; ++j) if (parent[j] >= 0) nchild[parent[j]]++;
        std::vector<int32_t> start;
        {
            std::vector<char> begins(n, 0);
            int32_t last = n - 1;
            int64_t zeros = 0;
            for (int32_t j = n - 1; j >= 1; --j) {
                bool join = false;
                if (parent[j - 1] == j) {
                    const int64_t cl = (int64_t)cs[last].size();
                    const int64_t c2 = last - j + 2, r2 = c2 + cl;
                    const int64_t zadd = (last - j + 1 + cl) - (int64_t)cs[j - 1].size();
                    const int64_t ents = c2 * r2 - c2 * (c2 + 1) / 2;
                    if (zadd == 0 && nchild[j] == 1) join = true;
                    else if (r2 <= AMALG_MAX_ORDER && c2 <= AMALG_MAX_PIVOTS && 5 * (zeros + zadd) <= ents) {
                        join = true;
                        zeros += zadd;
                    }
                }
                if (!join) { begins[j] = 1; last = j - 1; zeros = 0; }
            }
            start.push_back(0);
            for (int32_t j = 1; j < n; ++j) if (begins[j]) start.push_back(j);
        }
        start.push_back(n);
        const int32_t ns = (int32_t)start.size() - 1;
        t.col2sn.assign(n, 0);
        t.sn.assign(ns, Supernode{});
        for (int32_t sidx = 0; sidx < ns; ++sidx)
            for (int32_t j = start[sidx]; j < start[sidx + 1]; ++j) t.col2sn[j] = sidx;
        std::vector<std::vector<int32_t>> snkids(ns);
        int64_t u_off = 0, cb_off = 0, rcb_off = 0;
        t.row_idx.clear();
        for (int32_t sidx = 0; sidx < ns; ++sidx) {
            Supernode& S = t.sn[sidx];
            S.first = start[sidx];
            S.ncols = start[sidx + 1] - start[sidx];
            int32_t lastc = start[sidx + 1] - 1;
            S.nrows = (int32_t)cs[lastc].size();
            S.parent = parent[lastc] >= 0 ? t.col2sn[parent[lastc]] : -1;
            S.row_off = (int32_t)t.row_idx.size();
            t.row_idx.insert(t.row_idx.end(), cs[lastc].begin(), cs[lastc].end());
            const int64_t c = S.ncols, m = S.nrows, r = c + m;
            S.u_off = u_off; u_off += c * r;
            S.cb_off = cb_off; cb_off += m * m;
            S.rcb_off = rcb_off; rcb_off += m;
            for (int64_t k = 0; k < c; ++k) {
                int64_t rem = r - k - 1;
                t.lu_flops += 2 * rem * rem + rem;
            }
            t.max_front = std::max<int32_t>(t.max_front, (int32_t)r);
            if (S.parent >= 0) snkids[S.parent].push_back(sidx);
        }
        t.u_size = u_off; t.cb_size = cb_off; t.rcb_size = rcb_off;
        t.children.clear();
        for (int32_t sidx = 0; sidx < ns; ++sidx) {
            t.sn[sidx].child_begin = (int32_t)t.children.size();
            t.children.insert(t.children.end(), snkids[sidx].begin(), snkids[sidx].end());
            t.sn[sidx].child_end = (int32_t)t.children.size();
        }
        // levels (children have lower indices than their parents)
        int32_t nlev = 0;
        for (int32_t sidx = 0; sidx < ns; ++sidx) {
            int32_t l = 0;
            for (int32_t k = t.sn[sidx].child_begin; k < t.sn[sidx].child_end; ++k)
                l = std::max(l, t.sn[t.children[k]].level + 1);
            t.sn[sidx].level = l;
            nlev = std::max(nlev, l + 1);
        }
        t.n_levels = nlev;
        t.level_ptr.assign(nlev + 1, 0);
        for (auto& S : t.sn) t.level_ptr[S.level + 1]++;
        for (int32_t l = 0; l < nlev; ++l) t.level_ptr[l + 1] += t.level_ptr[l];
        t.level_sn.assign(ns, 0);
        {
            std::vector<int32_t> f(t.level_ptr.begin(), t.level_ptr.end());
            for (int32_t sidx = 0; sidx < ns; ++sidx) t.level_sn[f[t.sn[sidx].level]++] = sidx;
        }
        // front-local index of a permuted row x in supernode sidx
        auto local = [&](int32_t sidx, int32_t x) -> int32_t {
            const Supernode& S = t.sn[sidx];
            if (x >= S.first && x < S.first + S.ncols) return x - S.first;
            const int32_t* b = t.row_idx.data() + S.row_off;
            const int32_t* e = b + S.nrows;
            const int32_t* p = std::lower_bound(b, e, x);
            if (p == e || *p != x) return -1;
            return S.ncols + (int32_t)(p - b);
        };
        // relmap: each supernode's R rows -> local index in the parent's front
        t.relmap.assign(t.row_idx.size(), -1);
        for (int32_t sidx = 0; sidx < ns; ++sidx) {
            const Supernode& S = t.sn[sidx];
            if (S.parent < 0) {
                if (S.nrows) { err = "root supernode with update rows"; return NR_ERR_TOPOLOGY; }
                continue;
            }
            for (int32_t k = 0; k < S.nrows; ++k) {
                int32_t li = local(S.parent, t.row_idx[S.row_off + k]);
                if (li < 0) { err = "child update row outside the parent front"; return NR_ERR_TOPOLOGY; }
                t.relmap[S.row_off + k] = li;
            }
        }
        // A entries per supernode
        std::vector<std::vector<int32_t>> per(ns);
        for (int32_t r = 0; r < n; ++r)
            for (int32_t k = t.row_ptr[r]; k < t.row_ptr[r + 1]; ++k) {
                int32_t pr = t.iperm[r], pc = t.iperm[t.col_idx[k]];
                per[t.col2sn[std::min(pr, pc)]].push_back(k);
            }
        t.front_a_slot.clear(); t.front_a_li.clear(); t.front_a_lj.clear();
        for (int32_t sidx = 0; sidx < ns; ++sidx) {
            t.sn[sidx].a_begin = (int32_t)t.front_a_slot.size();
            for (int32_t k : per[sidx]) {
                int32_t r = -1;
                // row of CSR slot k
                r = (int32_t)(std::upper_bound(t.row_ptr.begin(), t.row_ptr.end(), k) - t.row_ptr.begin()) - 1;
                int32_t li = local(sidx, t.iperm[r]), lj = local(sidx, t.iperm[t.col_idx[k]]);
                if (li < 0 || lj < 0) { err = "matrix entry outside its front"; return NR_ERR_TOPOLOGY; }
                t.front_a_slot.push_back(k); t.front_a_li.push_back(li); t.front_a_lj.push_back(lj);
            }
            t.sn[sidx].a_end = (int32_t)t.front_a_slot.size();
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    t.analyse_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return NR_OK;
}

}  // namespace karam
