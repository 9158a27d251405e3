// seqgen.hpp -- the sequence generators of the reference's `config` and `dwi`
// subcommands: INI writers that prepare inputs for `sim`.
//
//   generate_config  src/config/config_generator.cpp:14-207 and
//                    src/config/handler.cpp:13-39 (GRE, SE, bSSFP templates written
//                    as a child config plus default_config.ini as its parent)
//   generate_pgse    src/dwi/pgse.cpp:15-153 (PGSE RF pair, gradient waveform and
//                    gradient scales inserted into an existing config)
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace swhost {

// seq_name is case-insensitive: "gre", "se" or "bssfp".  output is the child
// config path; default_config.ini is written beside it.
bool generate_config(const std::string &seq_name, uint32_t TE_us, uint32_t timestep_us,
                     const std::vector<std::string> &phantoms, const std::string &output, std::string &err);

// b_value[0] sets the gradient amplitude, every b_value a gradient scale
// sqrt(b / b_value[0]); dir is normalised; times in ms.
bool generate_pgse(const std::string &config_file, const std::vector<double> &b_value, std::vector<float> dir,
                   uint32_t start_ms, uint32_t delta_ms, uint32_t DELTA_ms, std::string &err);

}  // namespace swhost
