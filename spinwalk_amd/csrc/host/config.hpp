// config.hpp -- the sim config: INI reading with the reference's semantics and
// the parameter preparation that precedes a run.
#pragma once
#include <map>
#include <string>
#include <vector>

#include "../../../include/spinwalk_engine.h"

namespace swhost {

struct Config {
    sw_params p{};
    std::vector<double> diffusivity;
    std::vector<float> RF_FA, RF_PH, dephasing_deg, gx, gy, gz, pXY, T1, T2;
    std::vector<int32_t> TE, RF_t, dephasing_t, gradient_t;
    std::string seq_name, output_dir, top_config;
    std::map<std::string, std::vector<std::string>> files{{"PHANTOM", {}}, {"XYZ0", {}}, {"M0", {}}};
    std::vector<std::string> outputs;
    std::vector<float> scales;
    int32_t scale_type = 0;
    float linear_phase_cycling_orig = 0.f, quadratic_phase_cycling_orig = 0.f;

    Config();
    sw_seq seq_view() const;     // host views of the tables
};

// config_reader::prepare + parameters::prepare.  Returns false with a message.
bool load_config(const std::string &path, Config &c, std::string &err);

}  // namespace swhost
