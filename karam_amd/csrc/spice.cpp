// spice.cpp — SPICE netlist text -> the boundary netlist.
//
// Restates SPICEParser (netlist/spice_parser.cpp): the same line handling
// (title line, '*' comment lines, ';' inline comments, '+' continuation,
// .END), the same tokens ('(', ')', ',' and '=' are separators), the same
// value suffixes, the same node numbering (first appearance, 1-based; 0,
// GND, GROUND and VSS are ground) and element order, so a netlist parsed here
// has the node ids and component order the reference's parser gives it.
// Where that parser leaves the netlist undefined or drops what the line says,
// this one fills it in (each case is marked "fix" below):
//   - a diode's Is / N / Vt are never set there: here they come from its
//     .model card (IS, N), else 1e-14 A, 1 and 0.02585 V;
//   - a PULSE / SIN source with no DC value has an undefined DC value there:
//     here it is the waveform's value at t = 0 (v1, vo);
//   - SIN sources are typed "SIN", which the stamps (which look for "SINE")
//     never evaluate: here SIN (and SINE) select the sine waveform;
//   - the tokenizer splits "W=2u" into "W" "2u", so MOSFET W= / L= never
//     reach the device: here the name-value pairs are read;
//   - "AC mag [phase]" on a current source is read as DC values there: here it
//     is skipped as on a voltage source.
// Elements the batched Newton path does not simulate (L, K, Q, J, E, G, H, F,
// T, subcircuits) are refused with the line number rather than dropped.
#include "spice.h"

#include <cctype>
#include <map>
#include <sstream>
#include <stdexcept>

#include "karam_nr.h"

namespace karam {
namespace {

std::string upper(const std::string& s) {
    std::string r = s;
    for (auto& c : r) c = (char)std::toupper((unsigned char)c);
    return r;
}

// numeric prefix (std::stod) times the SI suffix; 0 when there is no number
// (parseValue, spice_parser.cpp:44-79)
double value_of(const std::string& s) {
    if (s.empty()) return 0.0;
    size_t used = 0;
    double v;
    try {
        v = std::stod(s, &used);
    } catch (...) {
        return 0.0;
    }
    if (used >= s.size()) return v;
    const std::string tail = s.substr(used);
    const std::string tu = upper(tail);
    if (tu == "MEG") return v * 1e6;
    if (tu == "MIL") return v * 25.4e-6;
    switch (std::tolower((unsigned char)tail[0])) {
    case 't': return v * 1e12;
    case 'g': return v * 1e9;
    case 'k': return v * 1e3;
    case 'm': return v * 1e-3;
    case 'u': return v * 1e-6;
    case 'n': return v * 1e-9;
    case 'p': return v * 1e-12;
    case 'f': return v * 1e-15;
    case 'a': return v * 1e-18;
    default: return v;
    }
}

std::vector<std::string> tokens_of(const std::string& line) {
    std::string flat = line;
    for (auto& c : flat)
        if (c == '(' || c == ')' || c == ',' || c == '=') c = ' ';
    std::vector<std::string> t;
    std::istringstream in(flat);
    std::string w;
    while (in >> w) t.push_back(w);
    return t;
}

struct Line {
    int number;
    std::string text;
};

// physical lines -> logical lines (preprocessLines, spice_parser.cpp:98-138)
std::vector<Line> logical_lines(const std::string& text) {
    std::vector<Line> out;
    std::istringstream in(text);
    std::string raw;
    int number = 0;
    while (std::getline(in, raw)) {
        ++number;
        const size_t semi = raw.find(';');
        if (semi != std::string::npos) raw.erase(semi);
        while (!raw.empty() && std::isspace((unsigned char)raw.back())) raw.pop_back();
        if (raw.empty() || raw[0] == '*') continue;
        if (raw[0] == '+' && !out.empty()) {
            out.back().text += " " + raw.substr(1);
            continue;
        }
        out.push_back({number, raw});
    }
    return out;
}

struct Card {
    std::string name, type;
    std::map<std::string, double> params;
};

class Parser {
public:
    Parser(int32_t flags, SpiceNetlist& o) : flags_(flags), o_(o) {}

    int run(const std::string& text, std::string& err) {
        const std::vector<Line> lines = logical_lines(text);
        if (lines.empty()) {
            err = "empty netlist";
            return NR_ERR_INVALID;
        }
        for (size_t k = 1; k < lines.size(); ++k) {     // lines[0] is the title
            const std::vector<std::string> t = tokens_of(lines[k].text);
            if (t.empty()) continue;
            o_.lines++;
            const std::string key = upper(t[0]);
            line_ = lines[k].number;
            if (key[0] == '.') {
                if (key == ".END") break;
                const int rc = directive(key, t, err);
                if (rc) return rc;
                continue;
            }
            const int rc = element(key[0], t, err);
            if (rc) return rc;
            o_.elements++;
        }
        finish();
        return NR_OK;
    }

private:
    int32_t node(const std::string& name) {
        const std::string u = upper(name);
        if (u == "0" || u == "GND" || u == "GROUND" || (u == "VSS" && !(flags_ & SPICE_VSS_IS_NET))) return 0;
        auto it = nodes_.find(u);
        if (it != nodes_.end()) return it->second;
        const int32_t id = (int32_t)o_.node_names.size() + 1;
        nodes_.emplace(u, id);
        o_.node_names.push_back(u);
        return id;
    }

    int refuse(const std::string& what, std::string& err) {
        err = "line " + std::to_string(line_) + ": " + what +
              " is not simulated by the batched Newton path (flatten or remove it)";
        return NR_ERR_UNSUPPORTED;
    }

    void warn(const std::string& what) { o_.warnings.push_back("line " + std::to_string(line_) + ": " + what); }

    int directive(const std::string& key, const std::vector<std::string>& t, std::string& err) {
        if (key == ".MODEL") {
            if (t.size() < 3) return NR_OK;
            Card c;
            c.name = t[1];
            c.type = upper(t[2]);
            for (size_t i = 3; i + 1 < t.size(); i += 2) c.params[upper(t[i])] = value_of(t[i + 1]);
            cards_[upper(c.name)] = c;           // a later card of the same name wins
        } else if (key == ".SUBCKT" || key == ".ENDS") {
            return refuse("a subcircuit definition", err);
        } else if (key == ".OP") {
            o_.has_op = 1;
        } else if (key == ".TRAN") {
            o_.has_tran = 1;
            if (t.size() >= 3) {
                o_.tran_step = value_of(t[1]);
                o_.tran_stop = value_of(t[2]);
                if (t.size() >= 4) o_.tran_start = value_of(t[3]);
            }
        } else if (key == ".AC") {
            o_.has_ac = 1;
        } else if (key == ".DC") {
            o_.has_dc = 1;
        }
        // .PARAM values are not substituted into element values (as in the
        // reference); .OPTION, .INCLUDE, .LIB, .GLOBAL are skipped
        return NR_OK;
    }

    // waveform parameters after the keyword at t[i]: up to `count` values,
    // the rest 0
    static void wave_params(const std::vector<std::string>& t, size_t i, int count, double* par) {
        for (int q = 0; q < 7; ++q) par[q] = 0.0;
        for (int q = 0; q < count && i < t.size(); ++q, ++i) par[q] = value_of(t[i]);
    }

    // V / I source body from t[3] on (parseVoltSrc / parseCurrSrc,
    // spice_parser.cpp:370-466): DC value, waveform kind and parameters
    void source(const std::vector<std::string>& t, bool voltage, double& dc, int32_t& wave, double* par,
                double& ac_mag, double& ac_phase) {
        bool dc_set = false;
        wave = NR_WAVE_DC;
        for (int q = 0; q < 7; ++q) par[q] = 0.0;
        ac_mag = ac_phase = 0.0;
        size_t i = 3;
        while (i < t.size()) {
            const std::string kw = upper(t[i]);
            if (kw == "DC" && i + 1 < t.size()) {
                dc = value_of(t[i + 1]);
                dc_set = true;
                wave = NR_WAVE_DC;
                i += 2;
            } else if (kw == "AC" && i + 1 < t.size()) {
                // the phase slot takes the next token whatever it is, as in
                // the reference (fix for current sources: AC is skipped, not
                // read as DC values)
                ac_mag = value_of(t[i + 1]);
                if (i + 2 < t.size()) {
                    ac_phase = value_of(t[i + 2]);
                    i += 3;
                } else {
                    i += 2;
                }
            } else if (kw == "PULSE") {
                wave = NR_WAVE_PULSE;
                wave_params(t, i + 1, 7, par);
                break;
            } else if (kw == "SIN" || kw == "SINE") {       // fix: "SINE" too, and evaluated
                wave = NR_WAVE_SINE;
                wave_params(t, i + 1, 6, par);
                break;
            } else {
                dc = value_of(t[i]);
                dc_set = true;
                wave = NR_WAVE_DC;
                ++i;
            }
        }
        (void)voltage;
        // fix: a waveform source without a DC value takes its t = 0 value
        if (!dc_set) dc = wave == NR_WAVE_PULSE ? par[0] : wave == NR_WAVE_SINE ? par[0] : 0.0;
    }

    int32_t model_ref(const std::string& name) {
        auto it = refs_.find(name);
        if (it != refs_.end()) return it->second;
        const int32_t id = (int32_t)ref_order_.size();
        refs_.emplace(name, id);
        ref_order_.push_back(name);
        return id;
    }

    int element(char type, const std::vector<std::string>& t, std::string& err) {
        switch (type) {
        case 'R':
            if (t.size() < 4) { warn("resistor with fewer than 3 fields skipped"); break; }
            o_.res_name.push_back(t[0]);
            o_.res_n1.push_back(node(t[1]));
            o_.res_n2.push_back(node(t[2]));
            o_.res_ohms.push_back(value_of(t[3]));
            break;
        case 'C':
            if (t.size() < 4) { warn("capacitor with fewer than 3 fields skipped"); break; }
            o_.cap_name.push_back(t[0]);
            o_.cap_n1.push_back(node(t[1]));
            o_.cap_n2.push_back(node(t[2]));
            o_.cap_farads.push_back(value_of(t[3]));
            break;
        case 'V':
        case 'I': {
            if (t.size() < 3) { warn("source with fewer than 2 nodes skipped"); break; }
            const bool v = type == 'V';
            const int32_t p = node(t[1]), q = node(t[2]);
            double dc = 0.0, par[7], mag, ph;
            int32_t wave;
            source(t, v, dc, wave, par, mag, ph);
            auto& name = v ? o_.vsrc_name : o_.isrc_name;
            auto& pos = v ? o_.vsrc_pos : o_.isrc_pos;
            auto& neg = v ? o_.vsrc_neg : o_.isrc_neg;
            auto& val = v ? o_.vsrc_volts : o_.isrc_amps;
            auto& wk = v ? o_.vsrc_wave : o_.isrc_wave;
            auto& wp = v ? o_.vsrc_wave_par : o_.isrc_wave_par;
            name.push_back(t[0]);
            pos.push_back(p);
            neg.push_back(q);
            val.push_back(dc);
            wk.push_back(wave);
            wp.insert(wp.end(), par, par + 7);
            if (v) {
                o_.vsrc_ac_mag.push_back(mag);
                o_.vsrc_ac_phase.push_back(ph);
            }
            break;
        }
        case 'D':
            if (t.size() < 4) { warn("diode with fewer than 3 fields skipped"); break; }
            o_.dio_name.push_back(t[0]);
            o_.dio_anode.push_back(node(t[1]));
            o_.dio_cathode.push_back(node(t[2]));
            dio_model_.push_back(t[3]);
            o_.dio_model.push_back(model_ref(t[3]));
            break;
        case 'M': {
            if (t.size() < 6) { warn("MOSFET with fewer than 5 fields skipped"); break; }
            o_.mos_name.push_back(t[0]);
            o_.mos_d.push_back(node(t[1]));
            o_.mos_g.push_back(node(t[2]));
            o_.mos_s.push_back(node(t[3]));
            o_.mos_b.push_back(node(t[4]));
            o_.mos_model.push_back(model_ref(t[5]));
            double w = 1e-6, l = 1e-6;                // Mosfet defaults (netlist/circuit.h:476-477)
            for (size_t i = 6; i + 1 < t.size();) {   // fix: name-value pairs
                const std::string k = upper(t[i]);
                if (k == "W") { w = value_of(t[i + 1]); i += 2; }
                else if (k == "L") { l = value_of(t[i + 1]); i += 2; }
                else if (k == "AD" || k == "AS" || k == "PD" || k == "PS" || k == "NRD" || k == "NRS" || k == "NF" ||
                         k == "M") i += 2;
                else ++i;
            }
            o_.mos_w.push_back(w);
            o_.mos_l.push_back(l);
            break;
        }
        case 'L': return refuse("an inductor", err);
        case 'K': return refuse("a mutual inductance", err);
        case 'Q': return refuse("a BJT", err);
        case 'J': return refuse("a JFET", err);
        case 'E': case 'G': case 'H': case 'F': return refuse("a controlled source", err);
        case 'T': return refuse("a transmission line", err);
        case 'X': return refuse("a subcircuit instance", err);
        default:
            warn("unknown element '" + t[0] + "' skipped");
            break;
        }
        return NR_OK;
    }

    // model entries: the cards (one per upper-case name, under the name its
    // last .model line gave, in the order of the netlist's card map: by exact
    // name, netlist/circuit.h:1078-1080), then each referenced name with no
    // card of exactly that name (tensorizeNetlist looks cards up by exact
    // name, tensors/physics_tensors.cpp:424-441)
    void finish() {
        o_.num_nodes = (int32_t)o_.node_names.size();
        std::map<std::string, const Card*> final_cards;
        for (const auto& kv : cards_) final_cards[kv.second.name] = &kv.second;
        std::map<std::string, int32_t> by_name;
        for (const auto& kv : final_cards) {
            const Card& c = *kv.second;
            by_name[c.name] = (int32_t)o_.model_name.size();
            o_.model_name.push_back(c.name);
            o_.model_type.push_back(c.type);
            o_.model_has_card.push_back(1);
            auto get = [&](const char* k, int32_t& has, double& v) {
                auto it = c.params.find(k);
                has = it != c.params.end();
                v = has ? it->second : 0.0;
            };
            int32_t h1, h2, h3;
            double v1, v2, v3;
            get("KP", h1, v1);
            get("VTO", h2, v2);
            get("LAMBDA", h3, v3);
            o_.has_kp.push_back(h1); o_.kp.push_back(v1);
            o_.has_vto.push_back(h2); o_.vto.push_back(v2);
            o_.has_lambda.push_back(h3); o_.lambda.push_back(v3);
        }
        o_.models = (int32_t)cards_.size();
        std::vector<int32_t> remap(ref_order_.size());
        for (size_t r = 0; r < ref_order_.size(); ++r) {
            auto it = by_name.find(ref_order_[r]);
            if (it != by_name.end()) { remap[r] = it->second; continue; }
            by_name[ref_order_[r]] = remap[r] = (int32_t)o_.model_name.size();
            o_.model_name.push_back(ref_order_[r]);
            o_.model_type.push_back("");
            o_.model_has_card.push_back(0);
            o_.has_kp.push_back(0); o_.kp.push_back(0.0);
            o_.has_vto.push_back(0); o_.vto.push_back(0.0);
            o_.has_lambda.push_back(0); o_.lambda.push_back(0.0);
        }
        for (auto& m : o_.mos_model) m = remap[m];
        for (auto& m : o_.dio_model) m = remap[m];
        // fix: diode parameters from the diode's card (found as the parser
        // finds cards, by upper-case name), else the model defaults
        // (models/diode.yaml: Is 1e-14 A, N 1, Vt 0.02585 V)
        for (const std::string& name : dio_model_) {
            double is = 1e-14, n = 1.0;
            auto it = cards_.find(upper(name));
            if (it != cards_.end()) {
                auto p = it->second.params.find("IS");
                if (p != it->second.params.end()) is = p->second;
                p = it->second.params.find("N");
                if (p != it->second.params.end()) n = p->second;
            }
            o_.dio_is.push_back(is);
            o_.dio_n.push_back(n);
            o_.dio_vt.push_back(0.02585);
        }
    }

    int32_t flags_;
    SpiceNetlist& o_;
    int line_ = 0;
    std::map<std::string, int32_t> nodes_;
    std::map<std::string, Card> cards_;          // upper-case name -> card
    std::map<std::string, int32_t> refs_;        // referenced model name -> first-reference id
    std::vector<std::string> ref_order_, dio_model_;
};

}  // namespace

int parse_spice(const std::string& text, int32_t flags, SpiceNetlist& out, std::string& err) {
    out = SpiceNetlist{};
    Parser p(flags, out);
    return p.run(text, err);
}

}  // namespace karam
