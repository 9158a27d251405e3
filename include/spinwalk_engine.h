/*
 * spinwalk_engine.h -- C ABI of the MI355X (gfx950) SpinWalk simulation engine.
 *
 * The drop-in boundary for the `spinwalk sim -c config.ini` path of the reference
 * (aghaeifar/SpinWalk v1.21.0).  The reference has no plugin API; its seams are
 *   - the kernel seam  cu_sim(param, param_uvec, pFieldMap, pMask, M0, XYZ0, M1, XYZ1, T)
 *                      (src/sim/kernels.cuh:24, src/sim/kernels.cu:55-63), launched by
 *                      monte_carlo::run per scale (src/sim/monte_carlo.cu:324-333);
 *   - the upload seam  "move to GPU memory" (src/sim/monte_carlo.cu:247-262) and
 *                      parameters_dvec::copy_from_host (src/sim/simulation_parameters.cuh:74-90);
 *   - the driver seam  monte_carlo::run(config) / sim::handler::execute
 *                      (src/sim/monte_carlo.cu:199-355, src/sim/handler.cu:8-17).
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions: plain C types only; every call returns an int status (SW_OK == 0,
 * negative on error) and the engine keeps a message for sw_last_error().  Device
 * pointers are plain `void*`/typed pointers into HBM; `stream` is a hipStream_t
 * passed as `void*` (NULL = the engine's own stream).  One engine per host thread.
 */
#ifndef SPINWALK_ENGINE_H
#define SPINWALK_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_API __attribute__((visibility("default")))

/* ---- status codes --------------------------------------------------------- */
#define SW_OK               0
#define SW_E_INVALID       -1   /* bad argument / inconsistent sizes                 */
#define SW_E_HIP           -2   /* HIP runtime error                                 */
#define SW_E_NOMEM         -3   /* device memory check failed (device_helper.cu:78-101) */
#define SW_E_REPLAY        -4   /* a replayed spin ran out of recorded draws         */
#define SW_E_CONFIG        -5   /* config file rejected (config_reader.cpp:195-324) */
#define SW_E_IO            -6   /* phantom/output file error                         */
#define SW_E_NODEVICE      -7   /* no usable gfx950 device                           */

/* ---- enums ------------------------------------------------------------------ */
#define SW_RNG_PHILOX       0   /* Philox4x32-10 keyed by (seed, global spin)        */
#define SW_RNG_REPLAY       1   /* per-spin draw streams recorded on the host        */

#define SW_PREC_F32         0   /* reference arithmetic: FP32 phase and M            */
#define SW_PREC_F32_KAHAN   1   /* FP32 M, Kahan-compensated FP32 phase              */
#define SW_PREC_F64         2   /* FP64 field, phase and M (the FP64 oracle path)    */

#define SW_FIELD_NONE       0
#define SW_FIELD_F32        1   /* tesla per 1 T, float32 (reference layout)         */
#define SW_FIELD_F16        2   /* float16 x per-phantom scale                       */
#define SW_FIELD_BF16       3   /* bfloat16 x per-phantom scale                      */
#define SW_FIELD_Q16        4   /* int16 fixed point x per-phantom scale (max|field| */
                                /* is 32767 steps); absolute error <= max|field|/65534 */

/* The step loop of a single production walk (sw_walk_opts.step_loop): one
 * attempt per trip, or two (the second formed from the first's position and
 * both gathers in flight together; the same draws, steps and results).  Auto
 * takes two when the gathered records exceed 256 MiB, the Infinity Cache of an
 * MI355X.  Replayed walks and batches take one. */
#define SW_STEP_LOOP_AUTO   0
#define SW_STEP_LOOP_ONE    1
#define SW_STEP_LOOP_TWO    2

#define SW_DTYPE_F32        0
#define SW_DTYPE_F64        1

typedef struct sw_engine   sw_engine;
typedef struct sw_phantom  sw_phantom;
typedef struct sw_sequence sw_sequence;

/* Mirror of `parameters` (src/sim/simulation_parameters.cuh:176-246), the POD the
 * reference passes by value to cu_sim, as it stands for one scale. */
typedef struct sw_params {
    double   fov[3];                  /* metres (FoV-scaled for WHAT_TO_SCALE=0)     */
    uint64_t phantom_size[3];
    int64_t  matrix_length;
    float    B0, c, s;                /* c, s = cos/sin(RF_FA[0]) from prepare()     */
    float    linear_phase_cycling, quadratic_phase_cycling;
    int32_t  timestep_us, TR_us, n_dummy_scan;
    uint32_t n_spins, n_timepoints, n_substrate, n_scales;
    uint64_t seed, max_iterations;
    int32_t  enCrossFOV, enRecordTrajectory, fieldmap_exist;
} sw_params;

/* Mirror of `parameters_hvec` (src/sim/simulation_parameters.cuh:39-65) as host
 * views.  Times are time-step indices (after config_reader::timing_scale) and
 * diffusivity is already converted by parameters::prepare (metres/axis/step). */
typedef struct sw_seq {
    const double  *diffusivity;                      /* [n_substrate]                */
    const float   *T1_ms, *T2_ms;                    /* [n_substrate]                */
    const float   *pXY;                              /* [n_substrate^2]              */
    uint32_t       n_substrate;
    const float   *RF_FA_deg, *RF_PH_deg; const int32_t *RF_t; uint32_t n_rf;
    const int32_t *TE_t; uint32_t n_te;
    const float   *dephasing_deg; const int32_t *dephasing_t; uint32_t n_dephasing;
    const float   *gradientX_mTm, *gradientY_mTm, *gradientZ_mTm;
    const int32_t *gradient_t; uint32_t n_gradient;
} sw_seq;

/* Per-launch options that the reference expresses through host-side copies. */
typedef struct sw_walk_opts {
    uint32_t spin_offset;      /* global index of local spin 0 (RNG key, dephasing)   */
    uint32_t n_local;          /* spins in this launch                                 */
    float    xyz_scale;        /* XYZ0 * s in float (monte_carlo.cu:278); 1 = none     */
    float    gradient_scale;   /* G * s in float (monte_carlo.cu:288-290); 1 = none    */
    int32_t  rng;              /* SW_RNG_*                                              */
    int32_t  precision;        /* SW_PREC_*; SW_PREC_F64 writes M1 as double          */
    int32_t  field_format;     /* which phantom copy to gather (SW_FIELD_*)            */
    int32_t  spatial_order;    /* 1 = visit spins in start-voxel brick order           */
    const float *replay_normals;  uint64_t replay_normal_stride;   /* device, [n_local][stride] */
    const float *replay_uniforms; uint64_t replay_uniform_stride;  /* device, [n_local][stride] */
    int32_t  step_loop;        /* SW_STEP_LOOP_*; 0 (auto) when zero-initialised       */
} sw_walk_opts;

/* Per-launch counters returned by the walk. */
typedef struct sw_walk_stats {
    uint64_t n_lost;            /* spins abandoned (out of range or stuck)             */
    uint64_t n_replay_exhausted;/* replayed spins that ran out of draws                */
    uint64_t n_steps;           /* accepted spin-steps (dummy scans included)          */
    uint64_t n_attempts;        /* attempted steps (accepted + rejected)               */
    double   kernel_ms;         /* walk-kernel time from HIP events                    */
} sw_walk_stats;

/* ---- engine ----------------------------------------------------------------- */
/* monte_carlo::monte_carlo(gpu_disabled, device_id) device selection
 * (src/sim/monte_carlo.cu:33-55) + check_CUDA (src/sim/device_helper.cu:18-30). */
SW_API int  sw_engine_create(int device_id, sw_engine **out);
SW_API void sw_engine_destroy(sw_engine *eng);
SW_API const char *sw_last_error(const sw_engine *eng);
/* print_device_info (src/sim/device_helper.cu:47-76): writes a text summary. */
SW_API int  sw_device_info(int device_id, char *buf, size_t buflen);
/* get_device_count (src/sim/device_helper.cu:32-45) */
SW_API int  sw_device_count(void);
SW_API const char *sw_version(void);

/* ---- host transforms -------------------------------------------------------- */
/* parameters::prepare (src/sim/simulation_parameters.cuh:227-245): c, s,
 * n_timepoints, diffusivity -> metres per axis per step (in place), automatic
 * dummy-scan count.  SEED == 0 draws from std::random_device as the reference does. */
SW_API int   sw_params_prepare(sw_params *p, double *diffusivity, float RF_FA0_deg, float T1_0_ms);
/* config_reader::timing_scale (src/sim/config_reader.cpp:39-46), in place. */
SW_API void  sw_timing_scale(int32_t *times_us, uint32_t n, int32_t timestep_us);
/* T -> degrees per step factor (src/sim/monte_carlo.cu:241). */
SW_API float sw_tesla2deg(float B0, int32_t timestep_us);
/* initialize_position, random branch (src/sim/monte_carlo.cu:141-151). */
SW_API void  sw_init_positions(uint64_t seed, const float fov_m[3], uint32_t n_spins, float *XYZ0);
/* The reference CPU path's draw streams for spins [spin_begin, spin_begin+n):
 * mt19937(seed+spin) discarded by seed+spin feeding normal_distribution<float>
 * and uniform_real_distribution<float> (src/sim/kernels.cu:82-88).  Row i holds
 * spin_begin+i.  n_threads host threads. */
SW_API int   sw_replay_fill(uint64_t seed, uint32_t spin_begin, uint32_t n,
                            float *normals, uint64_t normal_stride,
                            float *uniforms, uint64_t uniform_stride, int n_threads);

/* ---- phantom (read_phantom + fieldmap upload, monte_carlo.cu:98-123, 247-262) */
/* mask: uint8 [nx*ny*nz] row-major, z fastest.  field_T: float32 tesla per 1 T of
 * B0, same layout, or NULL.  formats: bitmask of (1<<SW_FIELD_F32) |
 * (1<<SW_FIELD_F16) | (1<<SW_FIELD_BF16) | (1<<SW_FIELD_Q16) copies to keep in HBM.  Host pointers;
 * the call is synchronous.  Fails with SW_E_NOMEM when the copies do not fit. */
SW_API int  sw_phantom_create(sw_engine *eng, const uint8_t *mask, const float *field_T,
                              const uint64_t dims[3], const float fov_m[3], uint32_t formats,
                              sw_phantom **out);
/* Same, from device-resident arrays (copied into engine-owned HBM). */
SW_API int  sw_phantom_create_device(sw_engine *eng, const uint8_t *d_mask, const float *d_field_T,
                                     const uint64_t dims[3], const float fov_m[3], uint32_t formats,
                                     sw_phantom **out);
SW_API void sw_phantom_destroy(sw_phantom *ph);
/* max(mask)+1 (src/sim/monte_carlo.cu:113) and the half-precision scale in use. */
SW_API int  sw_phantom_info(const sw_phantom *ph, uint32_t *n_labels, float *half_scale);

/* Copies a phantom back to the host (mask uint8 and field_T float32, either may
 * be NULL; field_T needs a float32 copy). */
SW_API int  sw_phantom_download(const sw_phantom *ph, uint8_t *mask, float *field_T);

/* ---- phantom generation on the device ---------------------------------------
 * Replaces the reference's host generators cylinder::run and sphere::run
 * (src/phantom/phantom_cylinder.cpp:85-301, src/phantom/phantom_sphere.cpp:79-198):
 * seeded non-overlapping cylinders (any axis) and spheres, labelled voxels and
 * the closed-form susceptibility field summed per voxel, B0 along z.  Specs are
 * applied in order (later labels win where shapes overlap).  achieved_vf
 * (nullable) receives the analytic volume fraction placed per spec, percent. */
#define SW_SHAPE_CYLINDER   0
#define SW_SHAPE_SPHERE     1
#define SW_SHAPE_AXONS      2   /* the reference cylinder generator's geometry: parallel
                                   to z, B0 tilted by orientation_deg in the x-z plane,
                                   optional random radii shrunk to fit (packed axons) */
typedef struct sw_shape_spec {
    int32_t  kind;                /* SW_SHAPE_*                                       */
    uint32_t label;               /* mask value inside the shape (1..255)             */
    float    radius_um;
    float    volume_fraction;     /* percent of the box                               */
    int32_t  random_orientation;  /* cylinders: isotropic axes                        */
    float    orientation_deg;     /* cylinders: axis tilt from B0 otherwise           */
    float    dchi, Y;             /* field scale 2pi/4pi * (1-Y) * dchi              */
    uint32_t seed;
    float    trunc_radii;         /* field truncation, in radii (0: cylinders none, spheres 10, axons 20) */
    float    radius_min_um;       /* axons: > 0, or radius_um < 0, draws radii in [radius_min, |radius|] */
} sw_shape_spec;
SW_API int  sw_phantom_generate(sw_engine *eng, const uint64_t dims[3], const float fov_um[3],
                                const sw_shape_spec *specs, uint32_t n_specs, uint32_t formats,
                                sw_phantom **out, double *achieved_vf);

/* ---- sequence tables (parameters_dvec::copy_from_host) ---------------------- */
SW_API int  sw_sequence_create(sw_engine *eng, const sw_params *p, const sw_seq *q, sw_sequence **out);
SW_API void sw_sequence_destroy(sw_sequence *seq);

/* ---- the kernel seam: cu_sim (src/sim/kernels.cu:55-63) ---------------------
 * Device pointers, asynchronous on `stream`.  M0, XYZ0: float [n_local][3];
 * M1: [n_local][n_TE][3] float (or double for SW_PREC_F64); T: uint8
 * [n_local][n_TE]; XYZ1: float [n_local][traj][3] with traj = (n_dummy+1)*
 * n_timepoints when enRecordTrajectory else 1.  Entries of M1/T of a spin that
 * is abandoned before an echo keep their prior contents (the reference's G9).
 * stats (nullable) is filled after an internal stream sync.  The counters,
 * visiting order and excitation tables are engine-owned: a call on another
 * stream than the previous call's waits on the device for that call's use of
 * them (so walks of one engine on several streams run one after the other). */
SW_API int  sw_walk_device(sw_engine *eng, const sw_phantom *ph, const sw_sequence *seq,
                           const sw_params *p, const sw_walk_opts *o,
                           const float *d_M0, const float *d_XYZ0,
                           void *d_M1, float *d_XYZ1, uint8_t *d_T,
                           void *stream, sw_walk_stats *stats);

/* sw_walk_device without allocation.  No counterpart in the reference: this
 * entry point is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as sw_walk_device.  The counters, visiting order and
 * sort buffers live in a fixed 32 MiB device array that the library holds on
 * every device it runs on until the process ends; with spatial_order the call
 * needs 256 + 4*n_local (rounded up to 256) bytes plus the spatial sort's
 * buffers (8*n_local plus 2 MiB for phantoms up to 512^3, 16 MiB up to
 * 1024^3), so n_local up to about 2.6 million spins (1.4 million on phantoms
 * above 512^3), and refuses larger walks with SW_E_INVALID before launching
 * anything (walks without spatial_order have no limit).  The sequence's excitation table must
 * match p (sw_sequence_create builds it for the parameters given there).
 * Every call on one device shares that device's array, including calls of
 * different engines: calls that overlap in time on one device race and must
 * be ordered by the caller.  Calls no hipMalloc or hipFree. */
SW_API int  sw_walk_device_noalloc(sw_engine *eng, const sw_phantom *ph, const sw_sequence *seq,
                                   const sw_params *p, const sw_walk_opts *o,
                                   const float *d_M0, const float *d_XYZ0,
                                   void *d_M1, float *d_XYZ1, uint8_t *d_T,
                                   void *stream, sw_walk_stats *stats);

/* Host-pointer convenience with the reference's semantics: M1/T/XYZ1 are
 * uploaded first (so untouched entries keep the caller's values), the walk runs,
 * results are copied back; synchronous. */
SW_API int  sw_walk_host(sw_engine *eng, const sw_phantom *ph, const sw_sequence *seq,
                         const sw_params *p, const sw_walk_opts *o,
                         const float *M0, const float *XYZ0,
                         void *M1, float *XYZ1, uint8_t *T,
                         const float *replay_normals, const float *replay_uniforms,
                         sw_walk_stats *stats);

/* ---- batched walk --------------------------------------------------------------
 * The reference walks one scale of one phantom per launch, synchronously
 * (src/sim/monte_carlo.cu:271-342).  A batch walks many such tiles (phantoms,
 * FoV/gradient/phase-cycling scales, sequences, spin ranges) in one launch: each
 * tile is exactly what one sw_walk_device call with the same arguments walks, and
 * gives the same results.  Every tile shares rng, precision and field_format (those
 * of items[0].opts).  sw_batch_create checks every tile, builds per-tile excitation
 * tables and visiting orders, deals the work over the 8 XCDs longest tile first,
 * and captures counter reset, walk and the optional per-tile FP64 echo sums as one
 * HIP graph.  The phantoms, sequences and device buffers named by the items must
 * outlive the batch.  No counterpart in the reference: it replaces the scale loop's
 * launches. */
typedef struct sw_batch sw_batch;
typedef struct sw_batch_item {
    const sw_phantom  *phantom;
    const sw_sequence *sequence;
    sw_params          params;        /* the tile's parameters, as for sw_walk_device   */
    sw_walk_opts       opts;          /* spin range, scales; replay streams per tile    */
    const float       *d_M0, *d_XYZ0;
    void              *d_M1;
    float             *d_XYZ1;
    uint8_t           *d_T;
    double            *d_echo_sums;   /* nullable: [n_TE][n_substrate][3] FP64 sums, as
                                         sw_echo_reduce_device writes them              */
} sw_batch_item;
SW_API int  sw_batch_create(sw_engine *eng, const sw_batch_item *items, uint32_t n_items, sw_batch **out);
/* Runs the batch on `stream` (NULL: the engine's).  per_item (nullable, n_items
 * entries) and total (nullable) are filled after a stream sync; kernel_ms is the
 * time of the whole batch. */
SW_API int  sw_batch_run(sw_batch *b, void *stream, sw_walk_stats *per_item, sw_walk_stats *total);
/* Tiles that walk the same trajectories -- the same phantom, seed, spins, start
 * buffers (M0, XYZ0 pointers), draw source, tissue tables and event times, with
 * only RF/excitation parameters, gradient amplitudes, dephasing and outputs
 * apart, as the phase-cycling and gradient scales of one config are -- are
 * walked together, up to four per walk, sharing draws, moves and gathers (FP32
 * and Kahan phases).  Every tile's outputs remain its separate walk's, bit for
 * bit.  *n_walks: walks of the batch; *n_fused_tiles: tiles inside shared walks.
 * No counterpart in the reference. */
SW_API int  sw_batch_info(const sw_batch *b, uint32_t *n_walks, uint32_t *n_fused_tiles);
/* Whether sw_batch_run launches the batch's captured HIP graph (*uses_graph = 1)
 * and the nodes of that graph (nullable outputs).  Capture failures are errors
 * of sw_batch_create / sw_batch_set_schedule / sw_batch_tune, so a batch that
 * was created runs from its graph.  Runs of one batch on different streams are
 * ordered one after the other.  No counterpart in the reference. */
SW_API int  sw_batch_graph_info(const sw_batch *b, int32_t *uses_graph, uint32_t *n_nodes);
SW_API void sw_batch_destroy(sw_batch *b);

/* ---- batch schedule and its autotuner (no counterpart in the reference) -----
 * How a batch is walked.  Its spins go in chunks of chunk_spins * spins_per_lane
 * (one workgroup of chunk_spins lanes each).  Every tile is split into 8 parts,
 * one per XCD; within an XCD, tiles of one cost class (a factor of two in steps
 * per spin) take turns in bands through their parts, longer classes first.
 * waves_per_simd caps the waves resident on a SIMD (through the workgroups' LDS
 * allocation), which sets how many spins share each XCD's L2; order_brick
 * is the edge of the voxel bricks whose Morton order is each tile's visiting
 * order.  A schedule changes only which spins run where and when: results are
 * the same bit for bit under every schedule. */
typedef struct sw_batch_schedule {
    uint32_t chunk_spins;    /* lanes per workgroup: 64, 128 or 256                    */
    uint32_t bands_shared;   /* bands per part when every tile of a class walks one phantom */
    uint32_t bands_mixed;    /* bands per part for the other classes                   */
    uint32_t spins_per_lane; /* 1 or 2: spins a lane walks, one after the other        */
    uint32_t waves_per_simd; /* 0: as many as registers allow; 2-8: at most that many   */
    uint32_t order_brick;    /* 4, 8, 16 or 32 voxels                                  */
} sw_batch_schedule;
SW_API int  sw_batch_get_schedule(const sw_batch *b, sw_batch_schedule *out);
/* Rebuilds the chunk table and the graph for `s` (a tuned schedule, say). */
SW_API int  sw_batch_set_schedule(sw_batch *b, const sw_batch_schedule *s);
/* Times full runs of the batch on `stream` under candidate schedules (chunk size,
 * the band counts of the class kinds present, spins per lane, resident waves
 * per SIMD, visiting-order brick, then every pair of chunk size and resident
 * waves, since the two interact), keeps the fastest and recaptures
 * the graph.  Each trial times a schedule as the faster of two full runs; at
 * most max_trials trials after the current schedule's (0: no limit), which
 * follows one untimed run.  *best_ms / *first_ms: the kept schedule's time and
 * the starting schedule's. */
SW_API int  sw_batch_tune(sw_batch *b, void *stream, uint32_t max_trials, sw_batch_schedule *best,
                          float *best_ms, float *first_ms);

/* ---- echo reduction (demo/spinwalk_bold.ipynb:733-736), FP64 ----------------
 * out[(te*n_sub + label)*3 + {0,1,2}] = {sum Mx, sum My, count} over spins with
 * T[spin][te] == label, summed in float64 in a fixed order (deterministic). */
SW_API int  sw_echo_reduce_device(sw_engine *eng, const void *d_M1, int m1_dtype, const uint8_t *d_T,
                                  uint32_t n_spins, uint32_t n_te, uint32_t n_sub,
                                  double *d_out, void *stream);

/* ---- utilities (no counterpart in the reference) ----------------------------
 * The production draws of spins [spin0, spin0+n_spins), attempts [0, n_att):
 * d_out[(i*n_att + k)*4 + {0,1,2,3}] = {x, y, z, u} (float, device).  Used to test
 * the generator's statistics; the walk draws exactly these values. */
SW_API int  sw_philox_draws_device(sw_engine *eng, uint64_t seed, uint32_t spin0, uint32_t n_spins,
                                   uint32_t n_att, float *d_out, void *stream);
/* The memory ceiling of the voxel gather: n_lanes lanes each read n_reads (a
 * multiple of 8) 8-byte words at uniformly random 128-byte lines of
 * d_buf[0, bytes) (bytes >= 128); d_sink[n_lanes] (uint64, device) receives a
 * checksum that keeps the reads live.  Synchronous; *ms = kernel time on the
 * stream.  Lines read = n_lanes * n_reads. */
SW_API int  sw_probe_gather_device(sw_engine *eng, const void *d_buf, uint64_t bytes, uint32_t n_lanes,
                                   uint32_t n_reads, uint64_t seed, uint64_t *d_sink, void *stream, float *ms);
/* Waits for `stream` (NULL: the engine's stream). */
SW_API int  sw_stream_sync(sw_engine *eng, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SPINWALK_ENGINE_H */
