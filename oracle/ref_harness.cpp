// ref_harness.cpp — thin C entry points over the reference engine's own
// compilable translation units (math/sparse_lu.cpp, math/linalg.h,
// netlist/spice_parser.cpp, solvers/ac_solver.h,
// infrastructure/factor_graph.cpp, infrastructure/treewidth_analyzer.cpp,
// tensors/physics_tensors.cpp, tensors/kernel_dispatch.cpp,
// physics/device_physics.h), compiled from /root/reference by oracle/Makefile
// into oracle/_ref/libkaram_ref.so.
//
// TEST INFRASTRUCTURE ONLY: it pins the oracle restatement against the
// reference code in the build container.  Nothing in the product links it.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "infrastructure/factor_graph.h"
#include "infrastructure/treewidth_analyzer.h"
#include "math/linalg.h"
#include "math/sparse_lu.h"
#include "netlist/circuit.h"
#include "netlist/spice_parser.h"
#include "solvers/ac_solver.h"
#include "solvers/convergence_engine.h"
#include "physics/device_physics.h"
#include "tensors/kernel_dispatch.h"
#include "tensors/physics_tensors.h"

#include "../include/karam_nr.h"

using acutesim::compute::KernelDispatcher;

namespace {

TensorNetlist make_netlist(const nr_netlist* nl) {
    TensorNetlist t;
    t.environment.ambient_temp_K = nl->ambient_temp_K;
    TensorBlock& g = t.globalBlock;
    for (int i = 0; i < nl->n_res; ++i) g.addResistor(nl->res_n1[i], nl->res_n2[i], nl->res_ohms[i]);
    for (int i = 0; i < nl->n_cap; ++i) g.addCapacitor(nl->cap_n1[i], nl->cap_n2[i], nl->cap_farads[i]);
    for (int i = 0; i < nl->n_vsrc; ++i) g.addVoltageSource(nl->vsrc_pos[i], nl->vsrc_neg[i], nl->vsrc_volts[i]);
    for (int i = 0; i < nl->n_isrc; ++i) g.addCurrentSource(nl->isrc_pos[i], nl->isrc_neg[i], nl->isrc_amps[i]);
    for (int i = 0; i < nl->n_diode; ++i) {
        g.addDiode("D" + std::to_string(i), nl->dio_anode[i], nl->dio_cathode[i], nl->dio_is[i], nl->dio_n[i],
                   nl->dio_vt[i]);
        int m = nl->dio_model ? nl->dio_model[i] : -1;
        if (m >= 0 && m < nl->n_models && nl->models[m].name) g.diodes.back().modelName = nl->models[m].name;
    }
    for (int i = 0; i < nl->n_mos; ++i) {
        int m = nl->mos_model ? nl->mos_model[i] : -1;
        std::string model = (m >= 0 && m < nl->n_models && nl->models[m].name) ? nl->models[m].name : "";
        g.addMosfet("M" + std::to_string(i), nl->mos_d[i], nl->mos_g[i], nl->mos_s[i], nl->mos_b[i], model,
                    nl->mos_w[i], nl->mos_l[i]);
    }
    for (int i = 0; i < nl->n_models; ++i) {
        const nr_model_card& c = nl->models[i];
        if (!c.type) continue;
        ModelCard mc;
        mc.name = c.name;
        mc.type = c.type;
        if (c.has_kp) mc.params["KP"] = c.kp;
        if (c.has_vto) mc.params["VTO"] = c.vto;
        if (c.has_lambda) mc.params["LAMBDA"] = c.lambda;
        t.addModelCard(mc);
    }
    t.numGlobalNodes = nl->num_nodes;
    return t;
}

}  // namespace

extern "C" {

// tensorizeNetlist (tensors/physics_tensors.cpp:305-493): SoA device arrays.
int ref_tensorize(const nr_netlist* nl, int32_t* d_a, int32_t* d_c, double* d_is, double* d_n, double* d_vt,
                  int32_t* m_d, int32_t* m_g, int32_t* m_s, int32_t* m_b, double* m_w, double* m_l,
                  double* m_kp, double* m_vth, double* m_lam, int32_t* m_pmos) {
    TensorNetlist t = make_netlist(nl);
    TensorizedBlock tb = tensorizeNetlist(t);
    if ((int)tb.diodes.size() != nl->n_diode || (int)tb.mosfets.size() != nl->n_mos) return 1;
    for (size_t i = 0; i < tb.diodes.size(); ++i) {
        d_a[i] = tb.diodes.node_a[i]; d_c[i] = tb.diodes.node_c[i];
        d_is[i] = tb.diodes.Is[i]; d_n[i] = tb.diodes.N[i]; d_vt[i] = tb.diodes.Vt[i];
    }
    for (size_t i = 0; i < tb.mosfets.size(); ++i) {
        m_d[i] = tb.mosfets.drains[i]; m_g[i] = tb.mosfets.gates[i];
        m_s[i] = tb.mosfets.sources[i]; m_b[i] = tb.mosfets.bodies[i];
        m_w[i] = tb.mosfets.W[i]; m_l[i] = tb.mosfets.L[i];
        m_kp[i] = tb.mosfets.Kp[i]; m_vth[i] = tb.mosfets.Vth[i]; m_lam[i] = tb.mosfets.lambda[i];
        m_pmos[i] = tb.mosfets.isPMOS[i] ? 1 : 0;
    }
    return 0;
}

// FactorGraph::fromTensorizedBlock + TreewidthAnalyzer::analyze, as wired by
// NetlistCompiler::compile (numCircuitNodes = numGlobalNodes).
int ref_min_fill(const nr_netlist* nl, int32_t* ordering, int64_t cap, int32_t* len, int32_t* treewidth) {
    TensorNetlist t = make_netlist(nl);
    TensorizedBlock tb = tensorizeNetlist(t);
    FactorGraph fg = FactorGraph::fromTensorizedBlock(tb, (uint32_t)nl->num_nodes);
    TreewidthAnalysis a = TreewidthAnalyzer::analyze(fg);
    if ((int64_t)a.eliminationOrdering.size() > cap) return 1;
    for (size_t i = 0; i < a.eliminationOrdering.size(); ++i) ordering[i] = (int32_t)a.eliminationOrdering[i];
    *len = (int32_t)a.eliminationOrdering.size();
    *treewidth = (int32_t)a.estimatedTreewidth;
    return 0;
}

// MatrixConstructor::buildPattern (math/linalg.h:281-431) over the given
// triplet coordinates; slot[k] = CachedCsrPattern::indexOf(row, col).
int ref_pattern(int32_t n, int32_t nst, const int32_t* rows, const int32_t* cols, int32_t* row_ptr,
                int32_t* col_idx, int64_t cap, int32_t* nnz, int32_t* slot) {
    MatrixConstructor mc;
    mc.clear();
    mc.setDimensions(n, n);
    for (int k = 0; k < nst; ++k) mc.add(rows[k], cols[k], 1.0);
    CachedCsrPattern p = mc.buildPattern();
    if ((int64_t)p.col_indices.size() > cap) return 1;
    for (int i = 0; i <= n; ++i) row_ptr[i] = p.row_pointer[i];
    for (size_t k = 0; k < p.col_indices.size(); ++k) col_idx[k] = p.col_indices[k];
    *nnz = p.nnz;
    for (int k = 0; k < nst; ++k) slot[k] = p.indexOf(rows[k], cols[k]);
    return 0;
}

// symbolicFactorize (math/sparse_lu.cpp:14-128) on a CSR pattern with a
// 0-based elimination ordering of length n.
int ref_symbolic(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const int32_t* ordering0,
                 int32_t* perm, int32_t* frow_ptr, int32_t* fcol_idx, int64_t cap, int32_t* nnz_f,
                 int32_t* fill_in) {
    CachedCsrPattern p;
    p.rows = n;
    p.cols = n;
    p.row_pointer.assign(row_ptr, row_ptr + n + 1);
    p.col_indices.assign(col_idx, col_idx + row_ptr[n]);
    p.nnz = row_ptr[n];
    p.buildPositionMap();
    std::vector<uint32_t> ord(ordering0, ordering0 + n);
    SymbolicFactorization s = symbolicFactorize(p, ord);
    if ((int64_t)s.filledPattern.col_indices.size() > cap) return 1;
    for (int i = 0; i < n; ++i) perm[i] = s.perm[i];
    for (int i = 0; i <= n; ++i) frow_ptr[i] = s.filledPattern.row_pointer[i];
    for (size_t k = 0; k < s.filledPattern.col_indices.size(); ++k) fcol_idx[k] = s.filledPattern.col_indices[k];
    *nnz_f = s.filledPattern.nnz;
    *fill_in = s.estimatedFillIn;
    return 0;
}

// KernelDispatcher-bound device kernels (generic, as the SIMD TUs are not
// compiled).
void ref_eval_diodes(int32_t count, const int32_t* a, const int32_t* c, const double* Is, const double* N,
                     const double* Vt, const double* v, int32_t nv, double* vd, double* id, double* gd) {
    DiodeTensor t;
    for (int i = 0; i < count; ++i) t.push(a[i], c[i], Is[i], N[i], Vt[i]);
    std::vector<double> volts(v, v + nv);
    KernelDispatcher::get().batch_diode_physics(t, volts);
    for (int i = 0; i < count; ++i) { vd[i] = t.v_d[i]; id[i] = t.i_d[i]; gd[i] = t.g_d[i]; }
}

void ref_eval_mosfets(int32_t count, const int32_t* d, const int32_t* g, const int32_t* s, const double* W,
                      const double* L, const double* Kp, const double* Vth, const double* lam,
                      const int32_t* pmos, const double* v, int32_t nv, double* vgs, double* vds, double* ids,
                      double* gm, double* gds) {
    MosfetTensor t;
    for (int i = 0; i < count; ++i) t.push(d[i], g[i], s[i], 0, W[i], L[i], Kp[i], Vth[i], lam[i], pmos[i] != 0);
    std::vector<double> volts(v, v + nv);
    KernelDispatcher::get().batch_mosfet_physics(t, volts);
    for (int i = 0; i < count; ++i) {
        vgs[i] = t.vgs[i]; vds[i] = t.vds[i]; ids[i] = t.ids[i]; gm[i] = t.gm[i]; gds[i] = t.gds[i];
    }
}

// solveLU_Pivoted (math/linalg.h:742-854); returns 0 ok, 2 singular.
int ref_lu_pivoted(int32_t n, const int32_t* row_ptr, const int32_t* col_idx, const double* vals,
                   const double* rhs, double* x) {
    Csr_matrix m;
    m.rows = n;
    m.cols = n;
    m.nnz = row_ptr[n];
    m.row_pointer.assign(row_ptr, row_ptr + n + 1);
    m.col_indices.assign(col_idx, col_idx + m.nnz);
    m.values.assign(vals, vals + m.nnz);
    std::vector<double> b(rhs, rhs + n);
    SolverResult r = solveLU_Pivoted(m, b);
    for (int i = 0; i < n; ++i) x[i] = r.solution[i];
    return r.converged ? 0 : 2;
}

double ref_pnjlim(double vnew, double vold, double vt, double vcrit) { return pnjlim(vnew, vold, vt, vcrit); }
double ref_vcrit(double vt, double is) { return compute_vcrit(vt, is); }
double ref_diode_i_clamped(double vd, double Is, double N, double Vt) {
    DiodeParams<double> p = {Is, N, Vt};
    return diode_current_clamped(vd, p, 30.0);
}

// SPICEParser::parse (netlist/spice_parser.cpp:148-261) of `text`, dumped
// as lines of text into out (cap bytes); returns the length needed.  Fields
// the parser leaves uninitialised are not read: a source's DC value only
// when its type is "DC", never a diode's Is / N / Vt.
int64_t ref_spice_dump(const char* text, char* out, int64_t cap) {
    acutesim::SPICEParser parser;
    acutesim::SPICEParseResult r = parser.parse(std::string(text));
    std::string o;
    char buf[512];
    auto put = [&](const char* fmt, auto... a) {
        std::snprintf(buf, sizeof(buf), fmt, a...);
        o += buf;
    };
    put("success %d\nerror_line %d\nlines %d\nelements %d\nmodels %d\nnodes %d\n", (int)r.success, r.errorLine,
        r.linesProcessed, r.elementsCreated, r.modelsCreated, r.netlist.numGlobalNodes);
    put("directives %d %d %d %d\ntran %.17g %.17g %.17g\n", (int)r.hasOP, (int)r.hasTRAN, (int)r.hasAC,
        (int)r.hasDC, r.tranStep, r.tranStop, r.tranStart);
    const TensorBlock& g = r.netlist.globalBlock;
    for (const auto& x : g.resistors)
        put("R %s %d %d %.17g\n", x.name.c_str(), x.nodeTerminal1, x.nodeTerminal2, x.resistance_ohms);
    for (const auto& x : g.capacitors)
        put("C %s %d %d %.17g\n", x.name.c_str(), x.nodePlate1, x.nodePlate2, x.capacitance_farads);
    for (const auto& x : g.voltageSources) {
        const bool dc = x.type == "DC";
        put("V %d %d %s %d %.17g %.17g %.17g", x.nodePositive, x.nodeNegative, x.type.empty() ? "-" : x.type.c_str(),
            (int)dc, dc ? x.voltage_V : 0.0, x.ac_mag, x.ac_phase);
        for (double q : {x.pulse_v1, x.pulse_v2, x.pulse_td, x.pulse_tr, x.pulse_tf, x.pulse_pw, x.pulse_per,
                         x.sine_vo, x.sine_va, x.sine_freq, x.sine_td, x.sine_theta, x.sine_phase})
            put(" %.17g", q);
        o += "\n";
    }
    for (const auto& x : g.currentSources) {
        const bool dc = x.type == "DC";
        put("I %d %d %s %d %.17g", x.nodePositive, x.nodeNegative, x.type.empty() ? "-" : x.type.c_str(), (int)dc,
            dc ? x.current_A : 0.0);
        for (double q : {x.pulse_v1, x.pulse_v2, x.pulse_td, x.pulse_tr, x.pulse_tf, x.pulse_pw, x.pulse_per,
                         x.sine_vo, x.sine_va, x.sine_freq, x.sine_td, x.sine_theta, x.sine_phase})
            put(" %.17g", q);
        o += "\n";
    }
    for (const auto& x : g.diodes)
        put("D %s %d %d %s\n", x.instanceName.c_str(), x.anode, x.cathode, x.modelName.c_str());
    for (const auto& x : g.mosfets)
        put("M %s %d %d %d %d %s %.17g %.17g\n", x.instanceName.c_str(), x.drain, x.gate, x.source, x.body,
            x.modelName.c_str(), x.w, x.l);
    for (const auto& kv : r.netlist.modelCards) {
        put("MODEL %s %s %d", kv.second.name.c_str(), kv.second.type.c_str(), kv.second.level);
        for (const auto& p : kv.second.params) put(" %s=%.17g", p.first.c_str(), p.second);
        o += "\n";
    }
    put("counts L %d K %d Q %d J %d E %d G %d H %d F %d T %d X %d S %d\n", (int)g.inductors.size(),
        (int)g.mutualInductors.size(), (int)g.bjts.size(), (int)g.jfets.size(),
        (int)g.voltageControlledVoltageSources.size(), (int)g.voltageControlledCurrentSources.size(),
        (int)g.currentControlledVoltageSources.size(), (int)g.currentControlledCurrentSources.size(),
        (int)g.transmissionLines.size(), (int)r.netlist.instances.size(), r.subcircuitsCreated);
    if ((int64_t)o.size() < cap) std::memcpy(out, o.c_str(), o.size() + 1);
    return (int64_t)o.size();
}

// ACSolver::solveAC (solvers/ac_solver.h:339-383, dense complex LU with
// partial pivoting) at the operating point v_op; out[f][node][2].
int ref_ac(const nr_netlist* nl, const double* v_op, double f_start, double f_stop, int32_t n_points,
           int32_t input_node, double amplitude, double* freqs, double* out) {
    TensorNetlist t = make_netlist(nl);
    std::vector<double> dc(v_op, v_op + nl->num_nodes);
    ACSolver::ACResult r = ACSolver::solveAC(t, dc, f_start, f_stop, n_points, input_node, amplitude);
    if ((int)r.frequencies.size() != n_points) return 1;
    for (int f = 0; f < n_points; ++f) {
        freqs[f] = r.frequencies[f];
        for (int i = 0; i < nl->num_nodes; ++i) {
            out[((size_t)f * nl->num_nodes + i) * 2] = r.nodeVoltages[f][i].real();
            out[((size_t)f * nl->num_nodes + i) * 2 + 1] = r.nodeVoltages[f][i].imag();
        }
    }
    return 0;
}

// ConvergenceEngine::classifyFailure on a history given as residuals and
// voltage deltas (the fields it reads)
int ref_classify_failure(const double* residual, const double* max_dv, int64_t n, int32_t singular) {
    std::vector<ConvergenceStep> h((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        h[(size_t)i].residualNorm = residual[i];
        h[(size_t)i].maxVoltageDelta = max_dv[i];
    }
    return (int)ConvergenceEngine::classifyFailure(h, singular != 0);
}

int ref_noise(const nr_netlist* nl, const double* v_op, int32_t output_node, double f_start, double f_stop,
              int32_t n_points, double temp_K, double* freqs, double* psd, double* vrms) {
    TensorNetlist t = make_netlist(nl);
    std::vector<double> dc(v_op, v_op + nl->num_nodes);
    ACSolver::NoiseResult r = ACSolver::solveNoise(t, dc, output_node, 0, f_start, f_stop, n_points, temp_K);
    if ((int)r.frequencies.size() != n_points) return 1;
    for (int f = 0; f < n_points; ++f) {
        freqs[f] = r.frequencies[f];
        psd[f] = r.output_noise_psd[f];
    }
    *vrms = r.integrated_noise_vrms;
    return 0;
}

}  // extern "C"
