// spice.h — SPICE netlist text -> the boundary netlist (host side).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace karam {

// The parsed netlist: every array in netlist (line) order, node ids 1-based
// (0 = ground), plus what the parse found.
struct SpiceNetlist {
    int32_t num_nodes = 0;
    std::vector<std::string> node_names;              // node k is node_names[k - 1]
    std::vector<std::string> res_name, cap_name, vsrc_name, isrc_name, dio_name, mos_name;
    std::vector<int32_t> res_n1, res_n2, cap_n1, cap_n2;
    std::vector<double> res_ohms, cap_farads;
    std::vector<int32_t> vsrc_pos, vsrc_neg, vsrc_wave, isrc_pos, isrc_neg, isrc_wave;
    std::vector<double> vsrc_volts, vsrc_wave_par, vsrc_ac_mag, vsrc_ac_phase;
    std::vector<double> isrc_amps, isrc_wave_par;
    std::vector<int32_t> dio_anode, dio_cathode, dio_model;
    std::vector<double> dio_is, dio_n, dio_vt;
    std::vector<int32_t> mos_d, mos_g, mos_s, mos_b, mos_model;
    std::vector<double> mos_w, mos_l;
    // model entries: `.model` cards (has_card = 1) then names referenced
    // without an exactly matching card
    std::vector<std::string> model_name, model_type;
    std::vector<int32_t> model_has_card, has_kp, has_vto, has_lambda;
    std::vector<double> kp, vto, lambda;
    // what the parse found
    int32_t lines = 0, elements = 0, models = 0;
    int32_t has_op = 0, has_tran = 0, has_ac = 0, has_dc = 0;
    double tran_step = 0.0, tran_stop = 0.0, tran_start = 0.0;
    std::vector<std::string> warnings;
};

enum { SPICE_VSS_IS_NET = 1 };

// Returns 0 (NR_OK), or an NR_ERR_* code with `err` naming the line.
int parse_spice(const std::string& text, int32_t flags, SpiceNetlist& out, std::string& err);

}  // namespace karam
