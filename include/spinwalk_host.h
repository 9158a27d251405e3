/*
 * spinwalk_host.h -- C ABI of the host-side driver (libspinwalk_host.so).
 *
 * The C++ host side of the `spinwalk sim -c config.ini` drop-in: the INI config
 * reader with the reference's semantics, HDF5 phantom input and result output in
 * the reference's layout, and the per-phantom / per-scale driver that calls the
 * engine of include/spinwalk_engine.h.
 */
#ifndef SPINWALK_HOST_H
#define SPINWALK_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "spinwalk_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sw_config sw_config;

/* config_reader::prepare = read (recursive PARENT_CONFIG) + check + timing_scale
 * (src/sim/config_reader.cpp:25-37, 61-324) followed by parameters::prepare
 * (src/sim/simulation_parameters.cuh:227-245).  Returns SW_E_CONFIG with a
 * message in *err (nullable, caller frees with sw_config_free_string) when the
 * file is rejected. */
SW_API int  sw_config_load(const char *path, sw_config **out, char **err);
SW_API void sw_config_destroy(sw_config *cfg);
SW_API void sw_config_free_string(char *s);
/* The prepared scalar parameters and table views (valid while cfg lives).  The
 * phantom-dependent fields of sw_params (fov, phantom_size, matrix_length) are
 * zero until a phantom is read. */
SW_API int  sw_config_params(const sw_config *cfg, sw_params *p, sw_seq *q);
/* SCALE[i] values and WHAT_TO_SCALE (src/sim/config_reader.cpp:150-157). */
SW_API int  sw_config_scales(const sw_config *cfg, const float **scales, uint32_t *n, int32_t *what_to_scale);
/* FILES section: key "PHANTOM", "XYZ0" or "M0", index i (absolute paths after
 * resolution), the output file of phantom i, and SEQ_NAME. */
SW_API const char *sw_config_file(const sw_config *cfg, const char *key, uint32_t i);
SW_API uint32_t    sw_config_n_files(const sw_config *cfg, const char *key);
SW_API const char *sw_config_output(const sw_config *cfg, uint32_t i);
SW_API const char *sw_config_seq_name(const sw_config *cfg);

/* monte_carlo::run (src/sim/monte_carlo.cu:199-355): every phantom of the
 * config, every scale, results written to OUTPUT_DIR/<SEQ_NAME>_<phantom>.h5 with
 * the reference's datasets M, XYZ, T, scales, TE. */
SW_API int  sw_run_config(sw_engine *eng, const char *config_path);
/* sw_run_config with the engine's modes: rng SW_RNG_*, precision SW_PREC_F32 or
 * SW_PREC_F32_KAHAN (M is stored as float32), field_format SW_FIELD_*, and
 * spatial_order 0/1. */
SW_API int  sw_run_config_ex(sw_engine *eng, const char *config_path, int32_t rng, int32_t precision,
                             int32_t field_format, int32_t spatial_order);
/* sw_run_config_ex with every option.  With RECORD_TRAJECTORY the spins of a
 * scale are walked in chunks of trajectory_chunk_spins (0: as many as fit in
 * 512 MiB of trajectory), through two device buffers in turn; each chunk's
 * trajectory is downloaded on a second stream while the next chunk walks.  The
 * results do not depend on the chunk size. */
typedef struct sw_run_options {
    int32_t  rng, precision, field_format, spatial_order;
    uint32_t trajectory_chunk_spins;
} sw_run_options;
SW_API int  sw_run_config_opts(sw_engine *eng, const char *config_path, const sw_run_options *opt);
/* sim::handler::execute (src/sim/handler.cu:8-17).  use_cpu != 0 is rejected with
 * SW_E_INVALID: this engine has no CPU path (the reference's `-p` switch). */
SW_API int  sw_sim_execute(int use_cpu, uint32_t device_id, const char *const *config_files,
                           uint32_t n_files);
/* HDF5 dataset access in the layout the reference reads and writes (h5_helper,
 * src/sim/h5_helper.h:48-129).  dtype: 0 float32, 1 uint8, 2 float64.  sw_h5_read
 * converts to dtype, returns the shape in dims[<=8]/ndims and, when buf is not NULL,
 * copies at most cap_elems elements.  sw_h5_write replaces an existing dataset. */
SW_API int  sw_h5_read(const char *file, const char *name, int32_t dtype, void *buf, uint64_t cap_elems,
                       uint64_t *dims, uint32_t *ndims);
SW_API int  sw_h5_write(const char *file, const char *name, int32_t dtype, const void *buf,
                        const uint64_t *dims, uint32_t ndims);
/* The `config` subcommand: config_generator::generate_gre/se/bssfp and write_ini
 * (src/config/config_generator.cpp:14-207, src/config/handler.cpp:13-39).  Writes
 * `output` (SEQ_NAME, PHANTOM[i], the sequence's scan parameters) and, beside it,
 * default_config.ini as its PARENT_CONFIG.  seq_name: "gre", "se" or "bssfp" in any
 * case; SW_E_CONFIG for another name. */
SW_API int  sw_generate_config(const char *seq_name, uint32_t TE_us, uint32_t timestep_us,
                               const char *const *phantoms, uint32_t n_phantoms, const char *output);
/* The `dwi` subcommand: pgse::run (src/dwi/pgse.cpp:67-153).  Rewrites config_file
 * in place with the PGSE RF pair, the gradient waveform along dir (normalised),
 * WHAT_TO_SCALE = 1 and SCALE[i] = sqrt(b_values[i] / b_values[0]).  The amplitude
 * comes from b_values[0], which must be positive (a zero there divides by zero in
 * the reference). */
SW_API int  sw_generate_pgse(const char *config_file, const double *b_values, uint32_t n_b, const float dir[3],
                             uint32_t start_ms, uint32_t delta_ms, uint32_t DELTA_ms);
/* Last error message of the host library on this thread. */
SW_API const char *sw_host_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SPINWALK_HOST_H */
