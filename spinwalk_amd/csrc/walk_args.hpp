// walk_args.hpp -- argument block of the walk kernel and the small per-event
// tables it reads.  Shared by the host side (engine.cpp, compiled by g++) and the
// device side (walk.hip, compiled by hipcc); plain C++ only.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SW_HD_INLINE __host__ __device__ __forceinline__
#else
#define SW_HD_INLINE inline
#endif

namespace sw {

// Which rotation xrot_withphase() reduces to for a given RF phase; the reference
// tests the phase against 0, 180, 90 and -90/270 exactly (src/sim/kernels.cuh:162-181).
enum RotKind : int32_t { ROT_X = 0, ROT_XNEG = 1, ROT_Y = 2, ROT_YNEG = 3, ROT_GENERAL = 4 };

// One RF pulse after the excitation (RF index >= 1).  Trigonometry is formed on
// the host exactly as the reference does it: FA through double sin/cos narrowed
// to float (kernels.cuh:203-206), the phase through sinf/cosf of a float
// argument (kernels.cuh:184-185).  *_d are the double versions for the FP64 path.
struct RfPulse {
    int32_t t;          // time index
    int32_t kind;       // RotKind for the float path
    int32_t kind_d;     // RotKind for the double path
    int32_t pad;
    float   s, c, sp, cp;
    double  sd, cd, spd, cpd;
};

// Excitation of dummy scan d: RF_FA[0] with phase RF_PH[0] + d*lin + d(d+1)/2*quad
// wrapped to [0, 360] (src/sim/kernels.cu:110-117).
struct Excitation {
    int32_t kind, kind_d;
    float   sp, cp;
    double  spd, cpd;
};

// Voxel records in 128-byte bricks (one cache line).  A record holds the
// substrate label and the field of one voxel, so a voxel change is one load,
// and a spin that moves to a neighbouring voxel usually stays in the same line.
//   Rec8: {float field_T, uint32 label}          bricks of 4 x 2 x 2 voxels
//   Rec4: {uint16 field bits, uint8 label, pad}   bricks of 4 x 4 x 2 voxels
// Within a brick voxels are row-major (z fastest); bricks are row-major too.
struct BrickShape { int lbx, lby, lbz; };
constexpr BrickShape kBrick8 = {2, 1, 1};
constexpr BrickShape kBrick4 = {2, 2, 1};

// Packed records of the production walk's float32 field: a 128-byte line holds
// a brick of 2 x 3 x 4 voxels as 24 records of 5 bytes, {label, float32 field}
// at bytes 5 j .. 5 j + 4 of the line for voxel (x, y, z) of the brick, slot
// j = (x * 3 + y) * 4 + z (bytes 120-127 are padding).  5 bytes a voxel instead
// of 8, so a line covers 24 voxels instead of 16; and one 8-byte load at byte
// 5 j (unaligned, within the line: it ends by byte 122) reads label and field
// together.  Bricks are row-major, as for BrickShape.
struct PackShape { int bx, by, bz, nvox, rec_bytes; };
constexpr PackShape kPackF32 = {2, 3, 4, 24, 5};
// The same bricks as structure-of-arrays records: the 24 labels at bytes 0-23
// of the line, then the 24 float32 fields, aligned, at bytes 24 + 4 j.  Two
// aligned loads a gather (label at byte j, field at byte 24 + 4 j) instead of
// one unaligned one.  As a PackShape its records are the 1-byte labels, so a
// record's byte offset is its label's.
constexpr PackShape kSoaF32 = {2, 3, 4, 24, 1};
// The 16-bit fields (float16, bfloat16, int16) pack the same way: 3-byte
// records {label, 16-bit field} in bricks of 3 x 3 x 4 voxels, 36 to a line
// (bytes 108-127 padding), read with one 4-byte load at byte 3 j (it ends by
// byte 107).  A spin leaves a 3 x 3 x 4 brick 8 % less often than a 4 x 4 x 2
// one (the sums of the inverse sides, 0.917 against 1.0).
constexpr PackShape kPack16 = {3, 3, 4, 36, 3};
// A brick fills its slots, and the load at the last record (8 bytes for the
// float32 records, 4 for the 16-bit ones) ends inside the 128-byte line.
static_assert(kPackF32.nvox == kPackF32.bx * kPackF32.by * kPackF32.bz &&
              (kPackF32.nvox - 1) * kPackF32.rec_bytes + 8 <= 128, "float32 records leave their line");
static_assert(kSoaF32.bx == kPackF32.bx && kSoaF32.by == kPackF32.by && kSoaF32.bz == kPackF32.bz &&
              kSoaF32.nvox == kPackF32.nvox && kSoaF32.nvox * (1 + 4) <= 128,
              "structure-of-arrays float32 bricks differ from the packed ones or leave their line");
static_assert(kPack16.nvox == kPack16.bx * kPack16.by * kPack16.bz &&
              (kPack16.nvox - 1) * kPack16.rec_bytes + 4 <= 128, "16-bit records leave their line");
SW_HD_INLINE uint64_t pack_brick_count(uint64_t n, int b) { return (n + b - 1) / b; }
// Packed arrays that the walk addresses in its 32-bit mode (below 4 GiB, with
// the per-axis tables in LDS) group their bricks in super-blocks of 4 x 4 x 2
// bricks, each a contiguous run of 32 lines (bricks row-major inside), so that
// the lines of neighbouring bricks sit together in memory; super-blocks are
// row-major.  Larger arrays keep single bricks (super-blocks of 1 x 1 x 1).
// Along y and z the count of super-blocks (nb_y, nb_z of WalkArgs) is rounded
// up to an odd number, so that super-blocks adjacent in y or x are an odd
// multiple of the super-block's size apart, not a multiple of a large power of
// two (the padding is never gathered).
struct SuperShape { int sx, sy, sz; };
constexpr SuperShape kSuper = {4, 4, 2};
constexpr int kSuperLines = kSuper.sx * kSuper.sy * kSuper.sz;
SW_HD_INLINE uint64_t pack_super_count(uint64_t n, int b, int s) { return (pack_brick_count(n, b) + s - 1) / s; }
SW_HD_INLINE uint64_t pack_pitch_y(uint64_t ny, const PackShape &S, bool sup) {
    return pack_super_count(ny, S.by, sup ? kSuper.sy : 1) | 1;
}
SW_HD_INLINE uint64_t pack_pitch_z(uint64_t nz, const PackShape &S, bool sup) {
    return pack_super_count(nz, S.bz, sup ? kSuper.sz : 1) | 1;
}
SW_HD_INLINE uint64_t pack_bytes(uint64_t nx, uint64_t ny, uint64_t nz, const PackShape &S, bool sup) {
    return pack_super_count(nx, S.bx, sup ? kSuper.sx : 1) * pack_pitch_y(ny, S, sup) * pack_pitch_z(nz, S, sup) *
           (sup ? kSuperLines : 1) * 128;
}
// Whether the packed array of an nx x ny x nz phantom is laid out in super-blocks.
SW_HD_INLINE bool pack_super(uint64_t nx, uint64_t ny, uint64_t nz, const PackShape &S) {
    return pack_bytes(nx, ny, nz, S, true) < (1ull << 32) && (nx + ny + nz) * 4 <= 65536;
}
// How a launch addresses the records: byte offsets from the per-axis LDS tables
// (32-bit mode), or computed per lookup in 64-bit arithmetic for arrays in single
// bricks, or for a batch whose items mix both layouts (the per-item flag
// super_blocks is read at each lookup).
enum AddrMode : int { ADDR_64 = 0, ADDR_32 = 1, ADDR_64_MIXED = 2 };
// Line of brick (bx, by, bz) in a packed array of pitch nb_y, nb_z.
SW_HD_INLINE uint64_t pack_line(uint64_t bx, uint64_t by, uint64_t bz, uint64_t nb_y, uint64_t nb_z, bool sup) {
    if (!sup) return (bx * nb_y + by) * nb_z + bz;
    const uint64_t sb = ((bx / kSuper.sx) * nb_y + by / kSuper.sy) * nb_z + bz / kSuper.sz;
    return sb * kSuperLines + ((bx % kSuper.sx) * kSuper.sy + by % kSuper.sy) * kSuper.sz + bz % kSuper.sz;
}

SW_HD_INLINE uint64_t brick_count(uint64_t n, int lb) { return (n + (1ull << lb) - 1) >> lb; }

// A fused walk carries up to kMaxViews views: tiles whose spins walk the same
// trajectories (the same draws, geometry, start positions, tissue tables and
// event times) and differ only in RF and excitation phases or flip angles,
// gradient amplitudes or dephasing, and in their output buffers.  View 0 is the
// WalkArgs fields themselves; views 1.. are FusedArgs::xv.  Each view's
// magnetisation and phase are formed by exactly the operations of its own walk.
constexpr int kMaxViews = 4;
struct WalkView {
    const RfPulse    *rf;
    const Excitation *exc;
    const float      *deph_deg;
    const float      *gx, *gy, *gz;
    float             exc_s, exc_c;
    double            exc_sd, exc_cd;
    float             gradient_scale;
    void             *M1;
    float            *XYZ1;
    uint8_t          *T;
    unsigned long long *counters;
};

struct WalkArgs {
    // ---- phantom -----------------------------------------------------------
    const void     *rec;          // bricked voxel records (Rec8, Rec4, packed or structure-of-arrays)
    uint32_t        nb_y, nb_z;   // bricks along y and z
    float           t2d;          // float T -> deg/step factor (monte_carlo.cu:241)
    float           h2d;          // half -> deg/step factor = half_scale * t2d
    double          t2d_f64;      // double T -> deg/step factor for the FP64 path
    int64_t         nx, ny, nz, matrix_length;
    double          nd[3];        // nx, ny, nz as doubles
    double          fov[3];
    double          s2g[3];       // phantom_size / fov (kernels.cu:98)
    // the production walk's voxel-unit positions (PosVox)
    float           s2gf[3];      // (float) s2g: metres -> voxels
    float           vsf[3];       // (float) fov / n: voxels -> metres
    float           nf[3];        // nx, ny, nz as floats: the FoV in voxels
    double          vs[3];        // fov / n in double (the FP64 walk's gradient phase)
    float           g2d;          // (float) 1e-3 dt 1e-6 gamma rad2deg: mT/m x m -> degrees

    // ---- sequence tables (device) -------------------------------------------
    const double   *diffusivity;  // [n_sub] metres per axis per step
    const float    *pXY;          // [n_sub^2]
    const float    *T1s, *T2s;    // [n_sub] seconds, (float)(ms * 1e-3) (kernels.cu:167-168)
    const double   *T1d, *T2d;    // [n_sub] seconds in double
    uint32_t        n_sub;
    uint32_t        tissue_lds;   // 1: each workgroup copies the tissue tables to LDS (tissue_lds_words)
    uint32_t        pairs;        // 1: a single production walk takes two attempts per trip of its step loop
    uint32_t        super_blocks; // 1: the packed records are grouped in super-blocks (kSuper)
    const RfPulse  *rf;           // [n_rf]; entry 0 unused (excitation)
    uint32_t        n_rf;
    const int32_t  *TE_t;         uint32_t n_te;
    const int32_t  *deph_t;       const float *deph_deg; uint32_t n_deph;
    const int32_t  *grad_t;       const float *gx, *gy, *gz; uint32_t n_grad;
    const Excitation *exc;        // [n_dummy + 1]
    float           exc_s, exc_c; // param.s, param.c
    double          exc_sd, exc_cd;

    // ---- scan ----------------------------------------------------------------
    uint32_t        n_timepoints, n_dummy;
    int32_t         timestep_us;
    uint32_t        n_spins_total;
    uint32_t        spin_offset, n_local;
    uint64_t        seed, max_iterations;
    int32_t         cross_fov, record_traj;
    float           xyz_scale, gradient_scale;
    uint64_t        traj_len;     // XYZ1 slots per spin

    // ---- spins ---------------------------------------------------------------
    const float    *M0, *XYZ0;
    void           *M1;
    float          *XYZ1;
    uint8_t        *T;
    const uint32_t *order;        // visiting order (local indices) or NULL

    // ---- replay --------------------------------------------------------------
    const float    *rn; uint64_t rn_stride;
    const float    *ru; uint64_t ru_stride;

    // ---- counters: [0] lost, [1] replay exhausted, [2] accepted steps, [3] attempts
    unsigned long long *counters;
};

// Production walks in the 32-bit mode with at most kTissueLdsMax tissues keep
// their per-tissue tables in LDS, after the per-axis tables: n_sub^2 crossing
// probabilities, T1, T2 and the three step sizes of each tissue, 32-bit words.
constexpr uint32_t kTissueLdsMax = 16;
SW_HD_INLINE uint32_t tissue_lds_words(const WalkArgs &a) {
    return a.tissue_lds ? a.n_sub * a.n_sub + 5 * a.n_sub : 0u;
}

// An item of the fused batched walk: view 0 is `a`, views 1..n_views-1 are xv;
// entries of xv beyond them repeat view 0 with counters = NULL, so the kernel
// walks kMaxViews views and the extra ones rewrite view 0's values.
struct FusedArgs {
    WalkArgs a;
    uint32_t n_views;
    WalkView xv[kMaxViews - 1];
};

// A work unit of the batched walk: positions [first, first + chunk) of one
// item's visiting order.  Padding slots carry item == kNoItem.
struct BatchChunk { uint32_t item, first; };
constexpr uint32_t kNoItem = 0xFFFFFFFFu;

}  // namespace sw
