// config.cpp -- reading and checking a sim config.
//
// Follows config_reader::read / check / timing_scale (src/sim/config_reader.cpp:
// 61-193, 195-324, 39-46) and parameters::prepare (src/sim/simulation_parameters.cuh:
// 227-245).  Numbers are converted exactly as there: TR and TIME_STEP through
// std::stof then truncated to int, time lists through an istream of floats cast to
// int32, NUMBER_OF_SPINS and MAX_ITERATIONS through std::stod, SEED and the
// switches through std::stoi, tissue values through std::stof, P_XY rows through
// an istream of doubles.  Differences, each deliberate: a value that does not
// parse is reported as a config error (the reference throws), a missing
// WHAT_TO_SCALE means 0 (the reference leaves it uninitialised), TIME_STEP must be
// positive (the reference divides by it), and every config starts from the
// defaults of `parameters` instead of inheriting the previous config's state.
#include "config.hpp"

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <iterator>
#include <random>
#include <sstream>

#include "ini.hpp"

namespace fs = std::filesystem;

namespace swhost {
namespace {

constexpr double kDeg2Rad = 0.0174532925199433;

template <class T>
std::vector<T> str2vec(const std::string &str) {          // config_reader.cpp:16-23
    std::istringstream ss(str);
    std::vector<T> v;
    for (std::istream_iterator<float> it(ss), end; it != end; ++it) v.push_back(static_cast<T>(*it));
    return v;
}

struct Reader {
    Config &c;
    std::string err;

    bool num(const std::string &what, const std::string &s, double &out, int kind) {
        try {
            if (kind == 0) out = std::stof(s);
            else if (kind == 1) out = std::stod(s);
            else out = std::stoi(s);
            return true;
        } catch (const std::exception &) {
            err = "cannot parse " + what + " = '" + s + "'";
            return false;
        }
    }

    bool read(const std::string &path) {
        if (!fs::exists(path)) { err = "Config-file does not exist: " + path; return false; }
        Ini ini;
        if (!ini.read(path)) { err = "Failed to read config file: " + path; return false; }
        const std::string parent = ini.get("GENERAL", "PARENT_CONFIG");          // :78-85
        if (!parent.empty()) {
            fs::path pp(parent);
            if (pp.is_relative()) pp = fs::absolute(path).parent_path() / pp;
            if (!read(pp.string())) return false;
        }
        const std::string sn = ini.get("GENERAL", "SEQ_NAME");                  // :87
        if (!sn.empty()) c.seq_name = sn;
        for (const char *key : {"PHANTOM", "XYZ0", "M0"}) {                     // :90-100
            auto &lst = c.files.at(key);
            if (ini.has("FILES", std::string(key) + "[0]")) lst.clear();
            for (int i = 0;; i++) {
                const std::string v = ini.get("FILES", std::string(key) + "[" + std::to_string(i) + "]");
                if (v.empty()) break;
                lst.push_back(v);
            }
            for (auto &p : lst)
                if (fs::path(p).is_relative())
                    p = fs::weakly_canonical(fs::absolute(path).parent_path() / p).string();
        }
        const std::string od = ini.get("FILES", "OUTPUT_DIR");                  // :102-104
        if (!od.empty()) c.output_dir = od;
        if (fs::path(c.output_dir).is_relative())   // against the top-level config, as the reference does
            c.output_dir = fs::weakly_canonical(fs::absolute(c.top_config).parent_path() / c.output_dir).string();

        double v;
        auto scan = [&](const char *k) { return ini.get("SCAN_PARAMETERS", k); };
        auto simp = [&](const char *k) { return ini.get("SIMULATION_PARAMETERS", k); };
        std::string s;
        if (!(s = scan("TR")).empty()) { if (!num("TR", s, v, 0)) return false; c.p.TR_us = (int32_t)(float)v; }
        if (!(s = scan("TIME_STEP")).empty()) { if (!num("TIME_STEP", s, v, 0)) return false; c.p.timestep_us = (int32_t)(float)v; }
        if (!(s = scan("TE")).empty()) c.TE = str2vec<int32_t>(s);
        if (!(s = scan("RF_T")).empty()) c.RF_t = str2vec<int32_t>(s);
        if (!(s = scan("RF_FA")).empty()) c.RF_FA = str2vec<float>(s);
        if (!(s = scan("RF_PH")).empty()) c.RF_PH = str2vec<float>(s);
        if (!(s = scan("DEPHASING_T")).empty()) c.dephasing_t = str2vec<int32_t>(s);
        if (!(s = scan("DEPHASING")).empty()) c.dephasing_deg = str2vec<float>(s);
        if (!(s = scan("GRADIENT_T")).empty()) c.gradient_t = str2vec<int32_t>(s);
        if (!(s = scan("GRADIENT_X")).empty()) c.gx = str2vec<float>(s);
        if (!(s = scan("GRADIENT_Y")).empty()) c.gy = str2vec<float>(s);
        if (!(s = scan("GRADIENT_Z")).empty()) c.gz = str2vec<float>(s);
        if (!(s = scan("DUMMY_SCAN")).empty()) { if (!num("DUMMY_SCAN", s, v, 2)) return false; c.p.n_dummy_scan = (int32_t)v; }
        if (!(s = scan("LINEAR_PHASE_CYCLING")).empty()) { if (!num("LINEAR_PHASE_CYCLING", s, v, 0)) return false; c.p.linear_phase_cycling = (float)v; }
        if (!(s = scan("QUADRATIC_PHASE_CYCLING")).empty()) { if (!num("QUADRATIC_PHASE_CYCLING", s, v, 0)) return false; c.p.quadratic_phase_cycling = (float)v; }
        if (!(s = simp("B0")).empty()) { if (!num("B0", s, v, 0)) return false; c.p.B0 = (float)v; }
        if (!(s = simp("SEED")).empty()) { if (!num("SEED", s, v, 2)) return false; c.p.seed = (uint64_t)(int64_t)(int32_t)v; }
        if (!(s = simp("NUMBER_OF_SPINS")).empty()) { if (!num("NUMBER_OF_SPINS", s, v, 1)) return false; c.p.n_spins = (uint32_t)v; }
        if (!(s = simp("CROSS_FOV")).empty()) { if (!num("CROSS_FOV", s, v, 2)) return false; c.p.enCrossFOV = (int32_t)v != 0; }
        if (!(s = simp("RECORD_TRAJECTORY")).empty()) { if (!num("RECORD_TRAJECTORY", s, v, 2)) return false; c.p.enRecordTrajectory = (int32_t)v != 0; }
        if (!(s = simp("MAX_ITERATIONS")).empty()) { if (!num("MAX_ITERATIONS", s, v, 1)) return false; c.p.max_iterations = (uint64_t)v; }
        if (ini.has("SIMULATION_PARAMETERS", "SCALE[0]")) c.scales.clear();     // :151-155
        for (int i = 0;; i++) {
            s = simp(("SCALE[" + std::to_string(i) + "]").c_str());
            if (s.empty()) break;
            if (!num("SCALE", s, v, 1)) return false;
            c.scales.push_back((float)v);
        }
        if (!(s = simp("WHAT_TO_SCALE")).empty()) { if (!num("WHAT_TO_SCALE", s, v, 2)) return false; c.scale_type = (int32_t)v; }
        auto tiss = [&](const std::string &k) { return ini.get("TISSUE_PARAMETERS", k); };
        std::vector<double> dif;                                                 // :160-166
        for (int i = 0; !(s = tiss("DIFFUSIVITY[" + std::to_string(i) + "]")).empty(); i++) {
            if (!num("DIFFUSIVITY", s, v, 0)) return false;
            dif.push_back((float)v);
        }
        if (!dif.empty()) c.diffusivity = dif;
        std::vector<float> t1, t2, pxy;
        for (int i = 0; !(s = tiss("T1[" + std::to_string(i) + "]")).empty(); i++) {
            if (!num("T1", s, v, 0)) return false;
            t1.push_back((float)v);
        }
        if (!t1.empty()) c.T1 = t1;
        for (int i = 0; !(s = tiss("T2[" + std::to_string(i) + "]")).empty(); i++) {
            if (!num("T2", s, v, 0)) return false;
            t2.push_back((float)v);
        }
        if (!t2.empty()) c.T2 = t2;
        for (int i = 0; !(s = tiss("P_XY[" + std::to_string(i) + "]")).empty(); i++) {   // :182-189
            std::istringstream iss(s);
            std::vector<double> row{std::istream_iterator<double>(iss), std::istream_iterator<double>()};
            for (double x : row) pxy.push_back((float)x);
        }
        if (!pxy.empty()) c.pXY = pxy;
        return true;
    }

    template <class T>
    bool ascending_unique(const std::vector<T> &v) {
        return std::is_sorted(v.begin(), v.end()) && std::adjacent_find(v.begin(), v.end()) == v.end();
    }

    bool check() {                                                              // :195-324
        auto fail = [&](const std::string &m) { err = m; return false; };
        if (c.RF_FA.size() != c.RF_t.size() || c.RF_FA.size() != c.RF_PH.size())
            return fail("RF_FA, RF_PH and RF_T must have the same number of elements");
        if (c.dephasing_t.size() != c.dephasing_deg.size())
            return fail("DEPHASING and DEPHASING_T must have the same number of elements");
        if (c.gx.size() != c.gy.size() || c.gx.size() != c.gz.size())
            return fail("GRADIENTS must have the same number of elements");
        if (c.gx.size() != c.gradient_t.size())
            return fail("GRADIENT_XYZ and GRADIENT_T must have the same number of elements");
        if (c.T1.size() != c.T2.size()) return fail("T1 and T2 must have the same number of elements");
        if (c.T1.size() != c.diffusivity.size()) return fail("T1 and diffusivity must have the same number of elements");
        if (c.T1.size() * c.T1.size() != c.pXY.size()) return fail("T1 and P_XY must have the same number of elements");
        if (c.scales.empty()) c.scales.push_back(1.0f);
        if (c.diffusivity.empty() || c.T1.empty() || c.T2.empty())
            return fail("Diffusivity, T1 and T2 must have at least one element");
        for (auto &kv : c.files)
            for (auto &p : kv.second)
                if (!fs::exists(p)) return fail("File does not exist: " + p);
        c.files.at("XYZ0").resize(c.files.at("PHANTOM").size(), "");
        c.files.at("M0").resize(c.files.at("PHANTOM").size(), "");
        try {
            fs::create_directories(fs::path(c.output_dir));
        } catch (const std::exception &e) {
            return fail("Creating directory " + c.output_dir + " failed. " + e.what());
        }
        c.outputs.clear();
        for (auto &p : c.files.at("PHANTOM")) {
            auto f = fs::path(c.output_dir) / (c.seq_name + "_" + fs::path(p).filename().string());
            c.outputs.push_back(f.replace_extension(".h5").string());
        }
        if (c.TE.empty() || !ascending_unique(c.TE) || c.TE[0] < 0)
            return fail("TE must exists and be in ascending order and must not have duplicates or negative values");
        if (c.RF_t.empty() || !ascending_unique(c.RF_t) || c.RF_t[0] != 0)
            return fail("RF times must be in ascending order, starts with 0 and must not have duplicates values");
        if (!ascending_unique(c.dephasing_t))
            return fail("Dephasing Times must be in ascending order and must not have duplicates values");
        if (!ascending_unique(c.gradient_t))
            return fail("Gradient times must be in a strickly ascending order and must not have duplicates values");
        if (c.p.TR_us < 0 || c.p.timestep_us <= 0) return fail("TR and timestep must be set");
        if (c.scale_type < 0 || c.scale_type > 2)
            return fail("WHAT_TO_SCALE must be 0, 1, or 2, but is " + std::to_string(c.scale_type));
        return true;
    }
};

}  // namespace

Config::Config() {
    // defaults of the reference's `parameters` (simulation_parameters.cuh:176-201)
    p.B0 = 9.4f;
    p.TR_us = -1;
    p.timestep_us = -1;
    p.n_dummy_scan = 0;
    p.linear_phase_cycling = 0.f;
    p.quadratic_phase_cycling = 0.f;
    p.n_spins = 1000;
    p.n_scales = 1;
    p.seed = 0;
    p.max_iterations = 9999;
    p.enCrossFOV = 1;
    p.enRecordTrajectory = 0;
    p.fieldmap_exist = 1;
}

sw_seq Config::seq_view() const {
    sw_seq q{};
    q.diffusivity = diffusivity.data(); q.n_substrate = (uint32_t)diffusivity.size();
    q.T1_ms = T1.data(); q.T2_ms = T2.data(); q.pXY = pXY.data();
    q.RF_FA_deg = RF_FA.data(); q.RF_PH_deg = RF_PH.data(); q.RF_t = RF_t.data(); q.n_rf = (uint32_t)RF_t.size();
    q.TE_t = TE.data(); q.n_te = (uint32_t)TE.size();
    q.dephasing_deg = dephasing_deg.data(); q.dephasing_t = dephasing_t.data(); q.n_dephasing = (uint32_t)dephasing_t.size();
    q.gradientX_mTm = gx.data(); q.gradientY_mTm = gy.data(); q.gradientZ_mTm = gz.data();
    q.gradient_t = gradient_t.data(); q.n_gradient = (uint32_t)gradient_t.size();
    return q;
}

bool load_config(const std::string &path, Config &c, std::string &err) {
    c = Config();
    c.top_config = path;
    Reader r{c, ""};
    if (!r.read(path) || !r.check()) { err = r.err; return false; }
    for (auto *v : {&c.TE, &c.RF_t, &c.dephasing_t, &c.gradient_t})       // timing_scale (:39-46)
        for (auto &x : *v) x = x / c.p.timestep_us;
    c.p.n_substrate = (uint32_t)c.diffusivity.size();
    c.p.n_scales = (uint32_t)c.scales.size();
    // parameters::prepare (simulation_parameters.cuh:227-245)
    if (sw_params_prepare(&c.p, c.diffusivity.data(), c.RF_FA[0], c.T1[0]) != SW_OK) {
        err = "parameters::prepare failed";
        return false;
    }
    c.linear_phase_cycling_orig = c.p.linear_phase_cycling;
    c.quadratic_phase_cycling_orig = c.p.quadratic_phase_cycling;
    (void)kDeg2Rad;
    return true;
}

}  // namespace swhost
