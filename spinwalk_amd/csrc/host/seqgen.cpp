// seqgen.cpp -- `config` and `dwi` sequence generators (see seqgen.hpp).
#include "seqgen.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <utility>

#include "ini.hpp"

namespace fs = std::filesystem;

namespace swhost {
namespace {

constexpr double kGamma = 267515315.0;   // rad/s/T (src/definitions.h:20)

std::string trimmed(std::string s) {
    const char *ws = " \t\n\r\f\v";
    s.erase(0, s.find_first_not_of(ws));
    s.erase(s.find_last_not_of(ws) + 1);
    return s;
}

// An INI file kept as its lines, so that values can be replaced in place with
// comments and layout preserved, as the reference's writer does when it updates
// an existing file (mINI INIFile::write).
class Doc {
public:
    bool load(const std::string &path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) return false;
        std::stringstream ss;
        ss << f.rdbuf();
        std::string text = ss.str();
        if (text.size() >= 3 && (unsigned char)text[0] == 0xEF && (unsigned char)text[1] == 0xBB &&
            (unsigned char)text[2] == 0xBF)
            text.erase(0, 3);
        lines_.clear();
        std::string cur;
        for (char c : text) {
            if (c == '\n') { lines_.push_back(cur); cur.clear(); }
            else if (c != '\r') cur.push_back(c);
        }
        if (!cur.empty()) lines_.push_back(cur);
        return true;
    }

    // Replace the last "key = ..." of the section, else add the key after the
    // section's last entry, else add the section at the end.
    void set(const std::string &sec, const std::string &key, const std::string &value) {
        const std::string line = key + " = " + value;
        std::string cur;
        long last_key = -1, last_entry = -1;
        for (size_t i = 0; i < lines_.size(); i++) {
            const std::string t = trimmed(lines_[i]);
            if (t.empty() || t[0] == ';') continue;
            if (t[0] == '[') {
                const size_t e = t.find(']');
                if (e != std::string::npos) {
                    cur = t.substr(1, e - 1);
                    if (cur == sec) last_entry = (long)i;
                }
                continue;
            }
            if (cur != sec) continue;
            last_entry = (long)i;
            const size_t eq = t.find('=');
            if (eq != std::string::npos && trimmed(t.substr(0, eq)) == key) last_key = (long)i;
        }
        if (last_key >= 0) { lines_[last_key] = line; return; }
        if (last_entry >= 0) { lines_.insert(lines_.begin() + last_entry + 1, line); return; }
        if (!lines_.empty() && !trimmed(lines_.back()).empty()) lines_.push_back("");
        lines_.push_back("[" + sec + "]");
        lines_.push_back(line);
    }

    bool save(const std::string &path) const {
        std::ofstream f(path, std::ios::binary | std::ios::trunc);
        if (!f) return false;
        for (size_t i = 0; i < lines_.size(); i++) f << lines_[i] << '\n';
        return (bool)f;
    }

private:
    std::vector<std::string> lines_;
};

// Sections and keys in insertion order, written fresh.
using Section = std::pair<std::string, std::vector<std::pair<std::string, std::string>>>;

class Fresh {
public:
    void add(const std::string &sec, const std::string &key, const std::string &value) {
        for (auto &s : secs_)
            if (s.first == sec) {
                for (auto &kv : s.second)
                    if (kv.first == key) { kv.second = value; return; }
                s.second.emplace_back(key, value);
                return;
            }
        secs_.push_back({sec, {{key, value}}});
    }
    bool save(const std::string &path) const {
        std::ofstream f(path, std::ios::binary | std::ios::trunc);
        if (!f) return false;
        for (size_t i = 0; i < secs_.size(); i++) {
            if (i) f << '\n';
            f << '[' << secs_[i].first << "]\n";
            for (auto &kv : secs_[i].second) f << kv.first << " = " << kv.second << '\n';
        }
        return (bool)f;
    }

private:
    std::vector<Section> secs_;
};

// The parent written beside every generated config
// (config_generator::generate_default_config, src/config/config_generator.cpp:14-87).
Fresh default_config(uint32_t TE_us, uint32_t timestep_us, const std::vector<std::string> &phantoms) {
    Fresh p;
    p.add("GENERAL", "PARENT_CONFIG", "");
    p.add("GENERAL", "SEQ_NAME", "noname");
    p.add("FILES", "OUTPUT_DIR", "./outputs");
    for (size_t i = 0; i < phantoms.size(); i++) p.add("FILES", "PHANTOM[" + std::to_string(i) + "]", phantoms[i]);
    p.add("FILES", "XYZ0[0]", "");
    p.add("FILES", "XYZ0[1]", "");
    p.add("FILES", "M0[0]", "");
    p.add("FILES", "M0[1]", "");
    p.add("TISSUE_PARAMETERS", "DIFFUSIVITY[0]", "1.0e-9");
    p.add("TISSUE_PARAMETERS", "DIFFUSIVITY[1]", "1.0e-9");
    p.add("TISSUE_PARAMETERS", "P_XY[0]", "1.0 0.0");
    p.add("TISSUE_PARAMETERS", "P_XY[1]", "0.0 1.0");
    p.add("TISSUE_PARAMETERS", "T1[0]", "2200");
    p.add("TISSUE_PARAMETERS", "T1[1]", "2200");
    p.add("TISSUE_PARAMETERS", "T2[0]", "41");
    p.add("TISSUE_PARAMETERS", "T2[1]", "41");
    p.add("SCAN_PARAMETERS", "TR", std::to_string(TE_us + timestep_us));
    p.add("SCAN_PARAMETERS", "TE", std::to_string(TE_us));
    p.add("SCAN_PARAMETERS", "RF_FA", "90.0");
    p.add("SCAN_PARAMETERS", "RF_PH", "0.0");
    p.add("SCAN_PARAMETERS", "RF_T", "0");
    for (const char *k : {"DEPHASING", "DEPHASING_T", "GRADIENT_X", "GRADIENT_Y", "GRADIENT_Z", "GRADIENT_T"})
        p.add("SCAN_PARAMETERS", k, "");
    p.add("SCAN_PARAMETERS", "TIME_STEP", std::to_string(timestep_us));
    p.add("SCAN_PARAMETERS", "DUMMY_SCAN", "0");
    p.add("SCAN_PARAMETERS", "LINEAR_PHASE_CYCLING", "0");
    p.add("SCAN_PARAMETERS", "QUADRATIC_PHASE_CYCLING", "0");
    p.add("SIMULATION_PARAMETERS", "B0", "9.4");
    p.add("SIMULATION_PARAMETERS", "SEED", "0");
    p.add("SIMULATION_PARAMETERS", "NUMBER_OF_SPINS", "1e5");
    p.add("SIMULATION_PARAMETERS", "CROSS_FOV", "0");
    p.add("SIMULATION_PARAMETERS", "RECORD_TRAJECTORY", "0");
    p.add("SIMULATION_PARAMETERS", "MAX_ITERATIONS", "1e4");
    p.add("SIMULATION_PARAMETERS", "WHAT_TO_SCALE", "0");
    p.add("SIMULATION_PARAMETERS", "SCALE[0]", "1.0");
    return p;
}

// A key's value through the PARENT_CONFIG chain, the child's value winning
// (dMRI::read, src/dwi/pgse.cpp:15-37).
bool read_chain(const std::string &file, const std::string &sec, const std::string &key, std::string &value,
                std::string &err, int depth = 0) {
    if (depth > 64) { err = "PARENT_CONFIG chain too deep at " + file; return false; }
    if (!fs::exists(file)) { err = "Config-file does not exist: " + file; return false; }
    Ini ini;
    if (!ini.read(file)) { err = "Failed to read config file: " + file; return false; }
    const std::string parent = ini.get("GENERAL", "PARENT_CONFIG");
    if (!parent.empty()) {
        fs::path pp(parent);
        if (pp.is_relative()) pp = fs::absolute(file).parent_path() / pp;
        if (!read_chain(pp.string(), sec, key, value, err, depth + 1)) return false;
    }
    const std::string v = ini.get(sec, key);
    if (!v.empty()) value = v;
    return true;
}

}  // namespace

bool generate_config(const std::string &seq_name, uint32_t TE_us, uint32_t timestep_us,
                     const std::vector<std::string> &phantoms, const std::string &output, std::string &err) {
    std::string seq = seq_name;
    std::transform(seq.begin(), seq.end(), seq.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (seq != "gre" && seq != "se" && seq != "bssfp") { err = "Invalid sequence name!"; return false; }
    const fs::path out = fs::weakly_canonical(fs::absolute(output));      // src/config/handler.cpp:15
    Fresh c;
    c.add("GENERAL", "SEQ_NAME", seq);
    for (size_t i = 0; i < phantoms.size(); i++) c.add("FILES", "PHANTOM[" + std::to_string(i) + "]", phantoms[i]);
    if (seq == "gre") {                                          // config_generator.cpp:90-117
        c.add("SCAN_PARAMETERS", "TR", std::to_string(TE_us + timestep_us));
        c.add("SCAN_PARAMETERS", "TE", std::to_string(TE_us));
        c.add("SCAN_PARAMETERS", "RF_FA", "90.0");
        c.add("SCAN_PARAMETERS", "RF_PH", "0");
        c.add("SCAN_PARAMETERS", "RF_T", "0");
        c.add("SCAN_PARAMETERS", "TIME_STEP", std::to_string(timestep_us));
    } else if (seq == "se") {                                    // :120-145
        c.add("SCAN_PARAMETERS", "TR", std::to_string(TE_us + timestep_us));
        c.add("SCAN_PARAMETERS", "TE", std::to_string(TE_us));
        c.add("SCAN_PARAMETERS", "RF_FA", "90.0 180.0");
        c.add("SCAN_PARAMETERS", "RF_PH", "0 90");
        c.add("SCAN_PARAMETERS", "RF_T", "0 " + std::to_string(TE_us / 2));
        c.add("SCAN_PARAMETERS", "TIME_STEP", std::to_string(timestep_us));
    } else {                                                     // :148-174
        c.add("SCAN_PARAMETERS", "TR", std::to_string(TE_us * 2));
        c.add("SCAN_PARAMETERS", "TE", std::to_string(TE_us));
        c.add("SCAN_PARAMETERS", "RF_FA", "16.0");
        c.add("SCAN_PARAMETERS", "RF_PH", "0");
        c.add("SCAN_PARAMETERS", "RF_T", "0");
        c.add("SCAN_PARAMETERS", "TIME_STEP", std::to_string(timestep_us));
        c.add("SCAN_PARAMETERS", "DUMMY_SCAN", "-1");
        c.add("SCAN_PARAMETERS", "LINEAR_PHASE_CYCLING", "180");
        c.add("SCAN_PARAMETERS", "QUADRATIC_PHASE_CYCLING", "0");
    }
    // write_ini (:177-200): the parent default_config.ini beside the child
    std::error_code ec;
    fs::create_directories(out.parent_path(), ec);
    if (ec) { err = "Creating directory " + out.parent_path().string() + " failed. " + ec.message(); return false; }
    const fs::path parent = out.parent_path() / "default_config.ini";
    c.add("GENERAL", "PARENT_CONFIG", parent.string());
    if (!c.save(out.string())) { err = "Failed to write config file: " + out.string(); return false; }
    if (!default_config(TE_us, timestep_us, phantoms).save(parent.string())) {
        err = "Failed to write config file: " + parent.string();
        return false;
    }
    return true;
}

bool generate_pgse(const std::string &config_file, const std::vector<double> &b_value, std::vector<float> dir,
                   uint32_t start_ms, uint32_t delta_ms, uint32_t DELTA_ms, std::string &err) {
    std::string ts;
    if (!read_chain(config_file, "SCAN_PARAMETERS", "TIME_STEP", ts, err)) return false;   // pgse.cpp:46-63
    if (ts.empty()) { err = "TIME_STEP is not set! Create a section SCAN_PARAMETERS and set TIME_STEP"; return false; }
    int timestep_us;
    try { timestep_us = std::stoi(ts); } catch (const std::exception &) { err = "cannot parse TIME_STEP: " + ts; return false; }
    if (timestep_us <= 0) { err = "TIME_STEP must be positive"; return false; }
    if (DELTA_ms < delta_ms) {                                    // :72-76
        err = "Delta must be greater than delta: " + std::to_string(DELTA_ms) + " vs " + std::to_string(delta_ms);
        return false;
    }
    if (b_value.empty() || dir.size() != 3) { err = "need b-values and a 3-component direction"; return false; }
    if (!(b_value[0] > 0)) { err = "the first b-value sets the gradient amplitude and must be positive"; return false; }
    const double d = delta_ms * 1e-3, D = DELTA_ms * 1e-3;         // :78-81
    const double G = std::sqrt(b_value[0] * 1e6 / (kGamma * kGamma * d * d * (D - d / 3.0))) * 1000.;
    const float n2 = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];   // :84-92, float sum
    const double norm = std::sqrt((double)n2);
    if (norm == 0) { err = "Direction vector is zero!"; return false; }
    for (auto &v : dir) v = (float)((double)v / norm);

    Doc doc;
    if (!doc.load(config_file)) { err = "Failed to read config file: " + config_file; return false; }
    const uint32_t start_us = start_ms * 1000;
    doc.set("SCAN_PARAMETERS", "RF_FA", "90 180");                  // :100-102
    doc.set("SCAN_PARAMETERS", "RF_PH", "0 90");
    doc.set("SCAN_PARAMETERS", "RF_T", "0 " + std::to_string(start_us + delta_ms * 1000 + (DELTA_ms - delta_ms) * 1000 / 2));

    const size_t n_points = delta_ms * 1000 / (uint32_t)timestep_us;      // :104-120
    std::string g[3] = {"0 ", "0 ", "0 "};
    for (int lobe = 0; lobe < 2; lobe++) {
        if (lobe) for (auto &s : g) s += "0 ";
        for (size_t i = 0; i < n_points; i++)
            for (int k = 0; k < 3; k++) g[k] += std::to_string(G * dir[k]) + " ";
        for (auto &s : g) s += "0 ";
    }
    doc.set("SCAN_PARAMETERS", "GRADIENT_X", g[0]);
    doc.set("SCAN_PARAMETERS", "GRADIENT_Y", g[1]);
    doc.set("SCAN_PARAMETERS", "GRADIENT_Z", g[2]);

    const uint64_t dt = (uint64_t)timestep_us, gap = (uint64_t)(DELTA_ms - delta_ms) * 1000;   // :126-135
    std::string t = std::to_string(start_us - (uint32_t)timestep_us) + " ";
    size_t i = 0;
    for (i = 0; i < n_points; i++) t += std::to_string(start_us + i * dt) + " ";
    t += std::to_string(start_us + i * dt + dt) + " ";
    t += std::to_string(start_us + gap + i * dt - dt) + " ";
    for (; i < 2 * n_points; i++) t += std::to_string(start_us + gap + i * dt) + " ";
    t += std::to_string(start_us + gap + i * dt + dt) + " ";
    doc.set("SCAN_PARAMETERS", "GRADIENT_T", t);

    doc.set("SIMULATION_PARAMETERS", "WHAT_TO_SCALE", "1");        // :139-142
    for (size_t k = 0; k < b_value.size(); k++)
        doc.set("SIMULATION_PARAMETERS", "SCALE[" + std::to_string(k) + "]",
                std::to_string(std::sqrt(b_value[k] / b_value[0])));
    if (!doc.save(config_file)) { err = "Failed to write config file: " + config_file; return false; }
    return true;
}

}  // namespace swhost
