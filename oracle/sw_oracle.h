/*
 * sw_oracle.h -- CPU oracle for the SpinWalk `sim` hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product (spinwalk_amd, include/) never
 * links, loads or calls anything under oracle/.
 *
 * It restates the reference's CPU path (C++ compiled without CUDA) line by
 * line, drawing through the same libstdc++ <random> objects in the same order,
 * so that on the same toolchain its draws and arithmetic are those of the
 * reference's `spinwalk sim -p` run.  Each function cites the reference lines
 * it follows (paths relative to the reference repository root).
 *
 * Parity pins (what fixes each function independently of this file):
 *   sub2ind, xrot, yrot, zrot, relax ... known answers of tests/test_kernel.cpp:17-89
 *   xrot_withphase ..................... axis-angle closed form (Rodrigues) and the
 *                                        exact special cases of kernels.cuh:162-181
 *   dephase_relax ...................... Bloch rotation + relaxation closed form
 *   sim() .............................. free precession, spin-echo magnitude,
 *                                        gradient phase, free-diffusion MSD,
 *                                        Stejskal-Tanner PGSE, impermeable and
 *                                        boundary invariants, brute force on tiny
 *                                        phantoms (tests/test_oracle_*.py)
 *   libstdc++ draw streams ............. mt19937 10000th output (C++ standard
 *                                        [rand.predef]) and an independent numpy
 *                                        restatement of the Marsaglia polar method
 *                                        over generate_canonical<float,24>
 *   prepare / timing_scale / Tesla2deg . closed forms (sqrt(2 D dt), integer
 *                                        division, B0*dt*gamma*180/pi)
 *   echo reduction ..................... brute force in numpy float64
 */
#ifndef SW_ORACLE_H
#define SW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of the reference's `parameters` POD (simulation_parameters.cuh:176-246),
 * values as they stand when the kernel is launched for one scale. */
typedef struct orc_params {
    double   fov[3];            /* metres, already FoV-scaled                     */
    uint64_t phantom_size[3];
    int64_t  matrix_length;
    float    c, s;              /* cos/sin of RF_FA[0] (prepare)                  */
    float    linear_phase_cycling, quadratic_phase_cycling;
    int32_t  timestep_us, TR_us, n_dummy_scan;
    uint32_t n_spins, n_timepoints, n_substrate;
    uint64_t seed, max_iterations;
    int32_t  enCrossFOV, enRecordTrajectory;
} orc_params;

/* Mirror of `parameters_uvec` (simulation_parameters.cuh:102-173): host views of
 * the sequence tables after config_reader::timing_scale (times are indices). */
typedef struct orc_tables {
    const double  *diffusivity;          /* [n_sub], metres per axis per step       */
    const float   *T1_ms, *T2_ms;        /* [n_sub]                                 */
    const float   *pXY;                  /* [n_sub*n_sub], row = from, col = to     */
    const float   *RF_FA_deg, *RF_PH_deg;
    const int32_t *RF_t;      uint32_t n_rf;
    const int32_t *TE_t;      uint32_t n_te;
    const float   *dephasing_deg;
    const int32_t *dephasing_t; uint32_t n_dephasing;
    const float   *gradientX_mTm, *gradientY_mTm, *gradientZ_mTm;
    const int32_t *gradient_t; uint32_t n_gradient;
} orc_tables;

/* ---- primitives (kernels.cuh) -------------------------------------------- */
int64_t orc_sub2ind3(int64_t x, int64_t y, int64_t z, int64_t nx, int64_t ny, int64_t nz);
int64_t orc_sub2ind4(int64_t x, int64_t y, int64_t z, int64_t o,
                     int64_t nx, int64_t ny, int64_t nz, int64_t no);
void orc_xrot_sc(float s, float c, const float *m0, float *m1);
void orc_yrot_sc(float s, float c, const float *m0, float *m1);
void orc_zrot_sc(float s, float c, const float *m0, float *m1);
void orc_xrot_deg(float theta, const float *m0, float *m1);
void orc_yrot_deg(float theta, const float *m0, float *m1);
void orc_zrot_deg(float theta, const float *m0, float *m1);
void orc_xrot_withphase_sc(float s, float c, float rf_phase, const float *m0, float *m1);
void orc_xrot_withphase_deg(float theta, float rf_phase, const float *m0, float *m1);
void orc_relax(float e1, float e2, const float *m0, float *m1);
void orc_dephase_relax(const float *m0, float *m1, float phase, float T1, float T2, float t);

/* ---- host transforms ----------------------------------------------------- */
/* parameters::prepare (simulation_parameters.cuh:227-245).  seed==0 is rejected
 * here (the reference draws std::random_device; an oracle must be reproducible).
 * diffusivity is converted in place; returns 0 on success. */
int orc_prepare(float RF_FA0_deg, int32_t TR_us, int32_t timestep_us, float T1_0_ms,
                double *diffusivity, uint32_t n_substrate, int32_t *n_dummy_scan,
                float *c, float *s, uint32_t *n_timepoints);
/* config_reader::timing_scale (config_reader.cpp:39-46) */
void orc_timing_scale(int32_t *v, uint32_t n, int32_t timestep_us);
/* Tesla -> degrees per time step (monte_carlo.cu:241) */
float orc_tesla2deg(float B0, int32_t timestep_us);
/* fieldmap *= factor in float (monte_carlo.cu:243-244) */
void orc_scale_field(float *field, uint64_t n, float factor);
/* initialize_position, random branch (monte_carlo.cu:141-151) */
void orc_init_positions(uint64_t seed, const float fov[3], uint32_t n_spins, float *XYZ0);
/* str2vec: istream_iterator<float> then static_cast (config_reader.cpp:16-23) */
int orc_str2vec_i32(const char *s, int32_t *out, int cap);

/* ---- the walk (kernels.cu:72-233) ----------------------------------------
 * Runs spins [spin_begin, spin_end) of one scale.  Output pointers are the
 * per-scale slices (monte_carlo.cu:318-320).  n_threads>1 spreads spins over
 * std::threads (the par_unseq of monte_carlo.cu:313).  Returns the number of
 * spins abandoned (out of range or stuck).  lost_flags (nullable) receives 1 for
 * each abandoned spin.  timing (nullable) receives {seconds in discard, seconds
 * total} summed over threads. */
uint32_t orc_sim(const orc_params *p, const orc_tables *q,
                 const float *field, const uint8_t *mask,
                 const float *M0, const float *XYZ0,
                 float *M1, float *XYZ1, uint8_t *T,
                 uint32_t spin_begin, uint32_t spin_end, int n_threads,
                 uint8_t *lost_flags, double *timing);

/* FP64 twin: identical trajectories and draws; field, phase, rotations, relaxation
 * and M carried in double.  field_T is the fieldmap in tesla per 1 T (un-scaled);
 * the T->deg/step factor is formed in double. */
uint32_t orc_sim_f64(const orc_params *p, const orc_tables *q, float B0,
                     const float *field_T, const uint8_t *mask,
                     const float *M0, const float *XYZ0,
                     double *M1, float *XYZ1, uint8_t *T,
                     uint32_t spin_begin, uint32_t spin_end, int n_threads,
                     uint8_t *lost_flags);

/* The draw streams sim() consumes for one spin: mt19937(seed+spin) discarded by
 * seed+spin, feeding normal_distribution<float>(0,1) (gen_r) and
 * uniform_real_distribution<float>(0,1) (gen_u) (kernels.cu:82-88). */
void orc_record_draws(uint64_t seed, uint32_t spin, float *normals, uint64_t n_normals,
                      float *uniforms, uint64_t n_uniforms);
/* Raw mt19937 outputs (for the standard's known answer). */
void orc_mt19937_raw(uint32_t seed, uint32_t *out, uint64_t n);

/* Echo reduction (demo/spinwalk_bold.ipynb:733-736) in float64:
 * out[(te*n_sub + label)*3 + {0,1,2}] = {sum Mx, sum My, count}. */
void orc_echo_reduce(const float *M1, const uint8_t *T, uint32_t n_spins, uint32_t n_te,
                     uint32_t n_sub, double *out);

#ifdef __cplusplus
}
#endif
#endif
