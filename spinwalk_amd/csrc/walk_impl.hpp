// walk_impl.hpp -- the per-spin Monte-Carlo walk for gfx950 (MI355X): device
// code and launch templates, instantiated by walk_replay.hip (replayed draws)
// and walk_philox.hip / walk_batch_philox.hip (production draws), so that the
// kernel families compile in parallel.
//
//
// One lane walks one spin through the whole sequence, as the reference kernel
// does (src/sim/kernels.cu:55-63 maps thread -> spin and runs sim()), but the
// state is laid out for CDNA4: positions stay in FP64 registers (full-rate FP64
// VALU on CDNA4, and FP64 is what the reference's trajectory arithmetic is,
// kernels.cu:93,131-136), event times are held as "next event" registers so a
// step costs four integer compares, per-event trigonometry is precomputed per
// sequence, and the per-step work is: draw -> move -> index -> (gather on voxel
// change) -> phase add.
//
// The step body follows src/sim/kernels.cu:72-233 expression for expression
// (float vs double promotion, truncation to int64 in sub2ind, the order of
// events within a step), so with replayed draws the trajectories, tissue labels
// and phases are the reference CPU path's bit for bit.  This file is compiled
// with -ffp-contract=off: a fused multiply-add would change roundings.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <math.h>
#include <type_traits>

#include "fmod360.hpp"
#include "philox.hpp"
#include "walk_args.hpp"
#include "walk_launch.hpp"

namespace sw {
namespace {

constexpr double kDeg2Rad = 0.0174532925199433;   // simulation_parameters.cuh:24
constexpr double kRad2Deg = 57.2957795130823;     // simulation_parameters.cuh:25
constexpr double kGamma   = 267515315.;           // definitions.h:20
constexpr uint32_t kNever = 0xFFFFFFFFu;
constexpr uint32_t kParked = 0x80000000u;   // set in a parked spin's time point (counts stay below 2^31)

enum : int { ST_OK = 0, ST_LOST = 1, ST_EXHAUSTED = 2 };



// ---------------------------------------------------------------------------
// Rotations (src/sim/kernels.cuh:82-152), generic over float/double.
// ---------------------------------------------------------------------------
template <class R>
__device__ __forceinline__ void rot_apply(int kind, R s, R c, R sp, R cp, R m[3]) {
    R a0 = m[0], a1 = m[1], a2 = m[2];
    switch (kind) {
    case ROT_X:    m[1] = c * a1 - s * a2;  m[2] = s * a1 + c * a2;  return;          // xrot(s,c)
    case ROT_XNEG: m[1] = c * a1 - (-s) * a2; m[2] = (-s) * a1 + c * a2; return;      // xrot(-s,c)
    case ROT_Y:    m[0] = c * a0 + s * a2;  m[2] = -s * a0 + c * a2; return;          // yrot(s,c)
    case ROT_YNEG: m[0] = c * a0 + (-s) * a2; m[2] = -(-s) * a0 + c * a2; return;     // yrot(-s,c)
    default: {
        // zrot(-sp,cp) -> xrot(s,c) -> zrot(sp,cp)   (kernels.cuh:183-188)
        const R nsp = -sp;
        R b0 = cp * a0 - nsp * a1, b1 = nsp * a0 + cp * a1, b2 = a2;
        R c0 = b0, c1 = c * b1 - s * b2, c2 = s * b1 + c * b2;
        m[0] = cp * c0 - sp * c1;
        m[1] = sp * c0 + cp * c1;
        m[2] = c2;
        return;
    }
    }
}

// dephase_relax (kernels.cu:45-52): zrot(phase deg) then relax(exp(-t/T1), exp(-t/T2)).
__device__ __forceinline__ void dephase_relax(float m[3], float phase, float T1, float T2, float t) {
    const float arg = (float)(phase * kDeg2Rad);            // float * double -> float argument
    const float s = sinf(arg), c = cosf(arg);
    const float a0 = m[0], a1 = m[1];
    m[0] = c * a0 - s * a1;
    m[1] = s * a0 + c * a1;
    if (T1 >= 0 && T2 >= 0) {
        // unqualified exp on a float argument resolves to double exp (kernels.cu:51)
        const float e1 = (float)exp((double)(-t / T1));
        const float e2 = (float)exp((double)(-t / T2));
        m[0] = m[0] * e2;                                    // relax (kernels.cuh:214-219)
        m[1] = m[1] * e2;
        m[2] = (float)(1. + e1 * (m[2] - 1.));
    }
}
__device__ __forceinline__ void dephase_relax(double m[3], double phase, double T1, double T2, double t) {
    const double arg = phase * kDeg2Rad;
    const double s = sin(arg), c = cos(arg);
    const double a0 = m[0], a1 = m[1];
    m[0] = c * a0 - s * a1;
    m[1] = s * a0 + c * a1;
    if (T1 >= 0 && T2 >= 0) {
        const double e1 = exp(-t / T1), e2 = exp(-t / T2);
        m[0] *= e2; m[1] *= e2; m[2] = 1. + e1 * (m[2] - 1.);
    }
}

// ---------------------------------------------------------------------------
// Phase accumulators.
// ---------------------------------------------------------------------------
struct PhaseF32 {          // reference: float accumulated_phase (kernels.cu:126,171)
    using R = float;
    float v = 0.f;
    __device__ void add(float x) { v += x; }
    __device__ void add_grad(double g) { v = (float)((double)v + g); }   // kernels.cu:185
    __device__ float value() const { return v; }
    __device__ float angle() const { return v; }     // what the rotation at an event uses
    __device__ void reset() { v = 0.f; }
};
struct PhaseKahan {        // FP32 with Kahan compensation
    using R = float;
    float v = 0.f, c = 0.f;
    __device__ void add(float x) {
        const float y = x - c;
        const float t = v + y;
        c = (t - v) - y;
        v = t;
    }
    __device__ void add_grad(double g) { add((float)g); }
    __device__ float value() const { return v - c; }
    // The rotation angle reduced modulo 360 before the compensation is applied
    // (an exact remainder, fmod360), so a phase of thousands of degrees keeps
    // its low bits.
    __device__ float angle() const { return fmod360(v) - c; }
    __device__ void reset() { v = 0.f; c = 0.f; }
};
struct PhaseF64 {
    using R = double;
    double v = 0.;
    __device__ void add(double x) { v += x; }
    __device__ void add_grad(double g) { v += g; }
    __device__ double value() const { return v; }
    __device__ double angle() const { return v; }
    __device__ void reset() { v = 0.; }
};

// ---------------------------------------------------------------------------
// Voxel records.  The reference multiplies the fieldmap by the T->deg/step factor
// in float once on the host (monte_carlo.cu:243-244); multiplying the stored
// field_T by the same float factor at the gather gives the same float.
// ---------------------------------------------------------------------------
// Loads through global (not flat) pointers: a flat load also counts against the
// LDS/scalar-load counter, so waiting on an LDS read or a scalar load would wait
// on the gather too; and with a 32-bit offset the load takes the base from
// scalar registers and the offset from one vector register.
template <class T> using GlobalPtr = const __attribute__((address_space(1))) T *;
template <class T, class I>
__device__ __forceinline__ T gload(const void *base, I byte_off) {
    const GlobalPtr<char> p = (GlobalPtr<char>)base + byte_off;
    if constexpr (std::is_same<T, uint2>::value) {
        const uint64_t v = *(GlobalPtr<uint64_t>)p;
        return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    } else {
        return *(GlobalPtr<T>)p;
    }
}

template <class T, bool IDX32>
__device__ __forceinline__ T load_rec(const void *base, uint64_t i) {
    if (IDX32) return gload<T>(base, (uint32_t)((uint32_t)i * (uint32_t)sizeof(T)));
    return gload<T>(base, i * sizeof(T));
}

struct RecF32 {                          // {float field_T, uint32 label}
    static constexpr BrickShape shape = kBrick8;
    static constexpr bool kPacked = false;
    static constexpr bool kMixedLayout = false;
    using T = uint2;
    __device__ static uint32_t label(T r) { return r.y; }
    template <class R> __device__ static R field(const WalkArgs &a, T r) {
        if (sizeof(R) == 8) return (R)((double)__uint_as_float(r.x) * a.t2d_f64);
        return (R)(__uint_as_float(r.x) * a.t2d);
    }
};
struct RecNone {                         // label only: {0, uint8 label, pad}
    static constexpr BrickShape shape = kBrick4;
    static constexpr bool kPacked = false;
    static constexpr bool kMixedLayout = false;
    using T = uint32_t;
    __device__ static uint32_t label(T r) { return r >> 16; }
    template <class R> __device__ static R field(const WalkArgs &, T) { return (R)0; }
};
// The 16-bit fields: packed 3-byte records {uint8 label, 16-bit field} in
// bricks of 3 x 3 x 4 voxels (kPack16).  A record index is the byte offset of
// the record; one 4-byte load there (unaligned, inside the line) brings the
// label (byte 0) and the field (bytes 1-2).
typedef uint32_t __attribute__((aligned(1))) u32_unaligned;
struct Rec16 {
    static constexpr bool kPacked = true;
    static constexpr bool kMixedLayout = false;
    static constexpr PackShape pack = kPack16;
    using T = uint32_t;                  // the 4 bytes loaded
    __device__ static uint32_t label(T r) { return r & 0xFFu; }
    __device__ static uint32_t bits(T r) { return (r >> 8) & 0xFFFFu; }
    template <bool IDX32, class I> __device__ static T load(const void *base, I off) {
        return gload<u32_unaligned>(base, off);
    }
};
struct RecF16 : Rec16 {                  // fp16 field_T / half_scale
    template <class R> __device__ static R field(const WalkArgs &a, T r) {
        return (R)(__half2float(__ushort_as_half((unsigned short)bits(r))) * a.h2d);
    }
};
struct RecQ16 : Rec16 {                  // int16 rint(field_T / q_scale)
    template <class R> __device__ static R field(const WalkArgs &a, T r) {
        return (R)((float)(int32_t)(int16_t)(uint16_t)bits(r) * a.h2d);
    }
};
struct RecBF16 : Rec16 {                 // bf16 field_T / half_scale
    template <class R> __device__ static R field(const WalkArgs &a, T r) {
        return (R)(__uint_as_float(bits(r) << 16) * a.h2d);
    }
};

// The production walk's float32 records: packed 5-byte records in bricks of
// 2x3x4 voxels, one 128-byte line each (kPackF32).  A record index is the byte
// offset of the record; one 8-byte load there (unaligned, inside the line)
// brings the label (byte 0) and the field (bytes 1-4).
typedef uint64_t __attribute__((aligned(1))) u64_unaligned;
struct RecF32P {
    static constexpr bool kPacked = true;
    static constexpr bool kMixedLayout = false;
    static constexpr PackShape pack = kPackF32;
    using T = uint2;                     // the 8 bytes loaded: x = bytes 0-3, y = bytes 4-7
    __device__ static uint32_t label(T r) { return r.x & 0xFFu; }
    template <class R> __device__ static R field(const WalkArgs &a, T r) {
        uint2 f;
        f.x = __builtin_amdgcn_alignbit(r.y, r.x, 8);                      // bytes 1-4
        return RecF32::field<R>(a, f);
    }
    template <bool IDX32, class I> __device__ static T load(const void *base, I off) {
        const uint64_t v = gload<u64_unaligned>(base, off);
        return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    }
};

// The single production walk's float32 records: the same bricks as
// structure-of-arrays (kSoaF32).  A record index is the byte offset of the
// label; the field sits 24 + 3 j bytes further, j = the slot in the brick =
// offset mod 128.  Two aligned loads from one line.
struct RecF32S {
    static constexpr bool kPacked = true;
    static constexpr bool kMixedLayout = false;
    static constexpr PackShape pack = kSoaF32;
    using T = uint2;                     // {field bits, label}, as RecF32
    __device__ static uint32_t label(T r) { return r.y; }
    template <class R> __device__ static R field(const WalkArgs &a, T r) { return RecF32::field<R>(a, r); }
    template <bool IDX32, class I> __device__ static T load(const void *base, I off) {
        const I foff = off + (I)3 * (off & (I)127);                       // field - nvox, as an offset
        T r;
        r.y = gload<uint8_t>(base, off);
        r.x = gload<uint32_t>((const char *)base + kSoaF32.nvox, foff);
        return r;
    }
};

// A packed record type whose arrays, in one batch, come in both layouts: the
// 64-bit lookups read each item's super_blocks flag (ADDR_64_MIXED).
template <class Rec> struct MixedLayout : Rec { static constexpr bool kMixedLayout = true; };

template <class Rec, bool IDX32, class I>
__device__ __forceinline__ typename Rec::T rec_load(const void *base, I off) {
    if constexpr (Rec::kPacked) return Rec::template load<IDX32>(base, off);
    else return load_rec<typename Rec::T, IDX32>(base, off);
}

// An empty statement that reads the next attempt's draws and claims to rewrite
// the gathered record: the draw arithmetic stays between the gather's issue
// and the first use of its result, hiding the load behind it.
__device__ __forceinline__ void after_draws(uint32_t &r, float x, float y, float z) {
    asm volatile("" : "+v"(r) : "v"(x), "v"(y), "v"(z));
}
__device__ __forceinline__ void after_draws(uint2 &r, float x, float y, float z) {
    asm volatile("" : "+v"(r.x), "+v"(r.y) : "v"(x), "v"(y), "v"(z));
}

// ---------------------------------------------------------------------------
// Draw sources.
// ---------------------------------------------------------------------------
// Both sources are used as peek/commit: peek() makes the normals of the next
// attempt available ahead of time (so they can be computed while the current
// attempt's gathers are in flight); commit() consumes them.  An attempt that does
// not move (diffusivity 0) consumes nothing, as in the reference (kernels.cu:130).
struct ReplayRng {          // the reference CPU path's recorded streams
    static constexpr bool kReplay = true;
    const float *n, *u;
    uint64_t nn, nu, in = 0, iu = 0;
    float px = 0.f, py = 0.f, pz = 0.f;
    bool avail = false;
    __device__ ReplayRng(const WalkArgs &a, uint32_t li, uint32_t) {
        n = a.rn + (uint64_t)li * a.rn_stride; nn = a.rn_stride;
        u = a.ru + (uint64_t)li * a.ru_stride; nu = a.ru_stride;
    }
    __device__ void peek() {
        avail = in + 3 <= nn;
        if (avail) { px = n[in]; py = n[in + 1]; pz = n[in + 2]; }
    }
    __device__ void commit() { in += 3; }
    __device__ bool uniform(float &v) {       // gen_u stream, one draw per substrate change
        if (iu >= nu) return false;
        v = u[iu++];
        return true;
    }
};

struct PhiloxRng {          // production: Philox4x32-10, one block per attempted step
    static constexpr bool kReplay = false;
    uint32_t k0, k1, m_hi, m_lo;      // key; M1 * spin, the first round's constant product
    uint64_t attempt = 0;
    float px = 0.f, py = 0.f, pz = 0.f;
    uint32_t pu = 0, ucur = 0;        // 24-bit permeability draws: next attempt's, last committed
    static constexpr bool avail = true;
    __device__ PhiloxRng(const WalkArgs &a, uint32_t, uint32_t spin_global)
        : k0((uint32_t)a.seed), k1((uint32_t)(a.seed >> 32)) {
        const uint64_t m = (uint64_t)kPhiloxM1 * spin_global;
        m_hi = (uint32_t)(m >> 32);
        m_lo = (uint32_t)m;
    }
    __device__ void peek() {
        const U32x4 r = philox4x32_10_pre((uint32_t)attempt, (uint32_t)(attempt >> 32), m_hi, m_lo, 0u, k0, k1);
        draw_bundle(r, px, py, pz, pu);
    }
    __device__ void commit() { ucur = pu; ++attempt; }
    // the permeability uniform of the attempt last committed
    __device__ bool uniform(float &v) { v = u24_to_unit(ucur); return true; }
    // The draws of attempt k (the bundle peek() makes when attempt == k).
    struct Draw { float x, y, z; uint32_t u; };
    __device__ Draw draw(uint64_t k) const {
        const U32x4 r = philox4x32_10_pre((uint32_t)k, (uint32_t)(k >> 32), m_hi, m_lo, 0u, k0, k1);
        Draw q;
        draw_bundle(r, q.x, q.y, q.z, q.u);
        return q;
    }
};

// ---------------------------------------------------------------------------
// sub2ind on truncated scaled positions (kernels.cu:140, kernels.cuh:53-60), and
// the record of that voxel in the bricked layout.
// When every scaled coordinate lies in [0, n) the truncations are the subscripts
// and the voxel is in range; its record index comes from per-axis tables in LDS
// (IDX32: every record byte offset < 2^32 and nx+ny+nz tables within 64 KiB) or
// from 64-bit arithmetic.  Anything else takes the reference's int64 arithmetic,
// including its behaviour for a subscript outside its axis (the index may still
// land inside the volume, in a neighbouring row); such a voxel is found again
// from the linear index.
// ---------------------------------------------------------------------------
// A voxel lookup: ok when the reference's row-major index lies in
// [0, matrix_length); rec is then the voxel's record index in the bricked
// layout, a bijection of the row-major index, so records compare as voxels do.
template <bool IDX32> using RecIdx = typename std::conditional<IDX32, uint32_t, uint64_t>::type;
template <bool IDX32> struct Vox {
    bool ok;
    RecIdx<IDX32> rec;
};

// The record index is separable: rec = X(ix) + Y(iy) + Z(iz) with
//   X = (bx * nb_y * nb_z) << L | (ix mod 2^lbx) << (lby + lbz)
//   Y = (by * nb_z) << L        | (iy mod 2^lby) << lbz
//   Z = bz << L                 |  iz mod 2^lbz
// (b = subscript >> lb, L = lbx + lby + lbz; the parts occupy disjoint bits, so
// the OR is a sum).  In the 32-bit mode each block keeps the three tables in LDS:
// axis_tab[0, nx) = X, [nx, nx+ny) = Y, [nx+ny, nx+ny+nz) = Z.
extern __shared__ uint32_t axis_tab[];

// For the packed bricks the index is the record's byte offset, separable in the
// same way (w bytes a record, brick coordinates b = i / brick side, super-blocks
// of sx x sy x sz bricks and SL = sx sy sz lines, pack_line):
// X = ((bx / sx) * nb_y * nb_z * SL + (bx mod sx) * sy * sz) * 128 + (ix mod bx) * by * bz * w,
// Y = ((by / sy) * nb_z * SL + (by mod sy) * sz) * 128 + (iy mod by) * bz * w,
// Z = ((bz / sz) * SL + bz mod sz) * 128 + (iz mod bz) * w.
// Subscripts, the in-super-block parts and the strides of a super-block along
// x (nb_y nb_z SL 128 bytes, checked below 2^32 on the host) and y fit 32 bits:
// the divisions run in 32-bit arithmetic and each part takes one 32 x 32 -> 64
// multiply-add in the 64-bit mode.
template <class Rec, int SX, int SY, int SZ, class I>
__device__ __forceinline__ I pack_part_of(const WalkArgs &a, int axis, I i_) {
    constexpr PackShape S = Rec::pack;
    constexpr uint32_t SL = SX * SY * SZ;
    const uint32_t i = (uint32_t)i_;
    if (axis == 0) {
        const uint32_t b = i / S.bx;
        const uint32_t stride = a.nb_y * a.nb_z * (SL * 128);
        return (I)(b / SX) * (I)stride +
               (I)((b % SX) * (uint32_t)(SY * SZ * 128) + (i % S.bx) * (uint32_t)(S.by * S.bz * S.rec_bytes));
    }
    if (axis == 1) {
        const uint32_t b = i / S.by;
        const uint32_t stride = a.nb_z * (SL * 128);
        return (I)(b / SY) * (I)stride + (I)((b % SY) * (uint32_t)(SZ * 128) + (i % S.by) * (uint32_t)(S.bz * S.rec_bytes));
    }
    const uint32_t b = i / S.bz;
    return (I)(b / SZ) * (I)(SL * 128) + (I)((b % SZ) * 128u + (i % S.bz) * (uint32_t)S.rec_bytes);
}

// The array's layout is uniform across the launch: a scalar branch picks the form.
template <class Rec, class I>
__device__ __forceinline__ I pack_part(const WalkArgs &a, int axis, I i) {
    if (a.super_blocks) return pack_part_of<Rec, kSuper.sx, kSuper.sy, kSuper.sz>(a, axis, i);
    return pack_part_of<Rec, 1, 1, 1>(a, axis, i);
}

template <class Rec>
__device__ __forceinline__ uint32_t axis_entry(const WalkArgs &a, uint32_t i) {
    const uint32_t nx = (uint32_t)a.nx, ny = (uint32_t)a.ny;
    if constexpr (Rec::kPacked) {
        if (i < nx) return pack_part<Rec, uint32_t>(a, 0, i);
        if (i < nx + ny) return pack_part<Rec, uint32_t>(a, 1, i - nx);
        return pack_part<Rec, uint32_t>(a, 2, i - nx - ny);
    } else {
        constexpr int LX = Rec::shape.lbx, LY = Rec::shape.lby, LZ = Rec::shape.lbz, L = LX + LY + LZ;
        if (i < nx)
            return (((i >> LX) * a.nb_y * a.nb_z) << L) | ((i & ((1u << LX) - 1)) << (LY + LZ));
        if (i < nx + ny) {
            i -= nx;
            return (((i >> LY) * a.nb_z) << L) | ((i & ((1u << LY) - 1)) << LZ);
        }
        i -= nx + ny;
        return ((i >> LZ) << L) | (i & ((1u << LZ) - 1));
    }
}

template <class Rec, bool IDX32>
__device__ __forceinline__ uint64_t record_of(const WalkArgs &a, uint64_t ix, uint64_t iy, uint64_t iz) {
    if (IDX32) {
        const uint32_t nx = (uint32_t)a.nx, nxy = nx + (uint32_t)a.ny;
        return axis_tab[(uint32_t)ix] + axis_tab[nx + (uint32_t)iy] + axis_tab[nxy + (uint32_t)iz];
    }
    if constexpr (Rec::kPacked) {
        // The 64-bit mode walks arrays in single bricks (an array in super-blocks
        // always fits the 32-bit mode), except in a batch that mixes layouts.
        if constexpr (Rec::kMixedLayout) {
            if (a.super_blocks)
                return pack_part_of<Rec, kSuper.sx, kSuper.sy, kSuper.sz>(a, 0, ix) +
                       pack_part_of<Rec, kSuper.sx, kSuper.sy, kSuper.sz>(a, 1, iy) +
                       pack_part_of<Rec, kSuper.sx, kSuper.sy, kSuper.sz>(a, 2, iz);
        }
        return pack_part_of<Rec, 1, 1, 1>(a, 0, ix) + pack_part_of<Rec, 1, 1, 1>(a, 1, iy) +
               pack_part_of<Rec, 1, 1, 1>(a, 2, iz);
    } else {
        constexpr int LX = Rec::shape.lbx, LY = Rec::shape.lby, LZ = Rec::shape.lbz;
        const uint64_t inner = ((ix & ((1u << LX) - 1)) << (LY + LZ)) | ((iy & ((1u << LY) - 1)) << LZ) |
                               (iz & ((1u << LZ) - 1));
        const uint64_t b = ((ix >> LX) * a.nb_y + (iy >> LY)) * a.nb_z + (iz >> LZ);
        return (b << (LX + LY + LZ)) | inner;
    }
}

// Truncating double -> int32 conversion with the instruction's defined
// saturation (a C cast out of range would be undefined).
__device__ __forceinline__ uint32_t cvt_i32_sat(double v) {
    uint32_t r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// A voxel some of whose scaled coordinates lie outside [0, n): the reference's
// int64 arithmetic on the truncated coordinates.
template <class Rec, bool IDX32>
__device__ __forceinline__ Vox<IDX32> locate_outside(const WalkArgs &a, double vx, double vy, double vz) {
    Vox<IDX32> v;
    v.ok = false;
    v.rec = 0;
    constexpr double kLim = 9.2e18;   // beyond int64 the conversion is undefined; treat as outside
    if (!(fabs(vx) < kLim && fabs(vy) < kLim && fabs(vz) < kLim)) return v;
    const int64_t jx = (int64_t)vx, jy = (int64_t)vy, jz = (int64_t)vz;
    const int64_t ind = jx * a.nz * a.ny + jy * a.nz + jz;     // kernels.cuh:53-60
    if (ind >= 0 && ind < a.matrix_length) {                    // kernels.cu:141
        const uint64_t u = (uint64_t)ind, nz = (uint64_t)a.nz, ny = (uint64_t)a.ny;
        v.ok = true;
        v.rec = (RecIdx<IDX32>)record_of<Rec, IDX32>(a, u / (nz * ny), (u / nz) % ny, u % nz);   // each in range
    }
    return v;
}

template <class Rec, bool IDX32>
__device__ __forceinline__ Vox<IDX32> locate(const WalkArgs &a, const double x[3]) {
    const double vx = x[0] * a.s2g[0], vy = x[1] * a.s2g[1], vz = x[2] * a.s2g[2];
    // v_cvt_i32_f64 truncates toward zero and saturates, so (unsigned)trunc(v) < n
    // holds exactly when the reference's int64 truncation lies in [0, n).  (It
    // maps NaN to 0; positions stay finite since every draw is finite.)
    const uint32_t ix = cvt_i32_sat(vx), iy = cvt_i32_sat(vy), iz = cvt_i32_sat(vz);
    if (ix < (uint32_t)a.nx && iy < (uint32_t)a.ny && iz < (uint32_t)a.nz) {
        Vox<IDX32> v;
        v.ok = true;
        v.rec = (RecIdx<IDX32>)record_of<Rec, IDX32>(a, ix, iy, iz);
        return v;
    }
    return locate_outside<Rec, IDX32>(a, vx, vy, vz);
}

// The boundary correction of one axis (kernels.cu:133-136) for v = xo + rw:
// v < 0: v + adj; v >= fov: v - adj, which rounds as v + (-adj); else v.
template <bool CROSS>
__device__ __forceinline__ double fix_axis(double fov, double rw, double v) {
    const double adj = CROSS ? fov : 2 * fabs(rw);
    const double sadj = v < 0 ? adj : -adj;
    return (v < 0 || v >= fov) ? v + sadj : v;
}

// v < 0 or v >= fov, or v is -0.0 (fov > 0 finite): the bit patterns of the
// non-negative doubles, read as unsigned integers, are ordered as the values,
// and every pattern with the sign bit set lies above them.  One 64-bit compare
// instead of two FP64 compares; fix_axis makes the exact decision.
__device__ __forceinline__ uint32_t maybe_outside(double v, double fov) {
    return (uint64_t)__double_as_longlong(v) >= (uint64_t)__double_as_longlong(fov) ? 1u : 0u;
}

// A step of all three axes (kernels.cu:130-137).  Spins rarely reach a wall, so
// the correction runs only when some lane of the wave may need it.
template <bool CROSS>
__device__ __forceinline__ void move(const WalkArgs &a, float px, float py, float pz, double diff, const double xo[3],
                                     double xn[3]) {
    const double rw0 = (double)px * diff, rw1 = (double)py * diff, rw2 = (double)pz * diff;   // :131
    xn[0] = xo[0] + rw0;
    xn[1] = xo[1] + rw1;
    xn[2] = xo[2] + rw2;
    const bool out = (maybe_outside(xn[0], a.fov[0]) | maybe_outside(xn[1], a.fov[1]) |
                      maybe_outside(xn[2], a.fov[2])) != 0;
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(out != 0) != 0, 0)) {
        xn[0] = fix_axis<CROSS>(a.fov[0], rw0, xn[0]);
        xn[1] = fix_axis<CROSS>(a.fov[1], rw1, xn[1]);
        xn[2] = fix_axis<CROSS>(a.fov[2], rw2, xn[2]);
    }
}

// ---------------------------------------------------------------------------
// Position arithmetic.  PosRef is the reference's: FP64 metres, its boundary
// correction and its sub2ind (kernels.cu:93-102, 130-140), so that replayed
// draws give the reference's trajectories bit for bit.  PosVox is the
// production walk's: FP32 positions in voxel units, so a step is one fused
// multiply-add per axis and a voxel subscript one truncation, with the FoV
// [0, n) per axis.  Both take the same draws and the same step sizes; PosVox
// rounds differently, so its walks agree with the reference in distribution.
// ---------------------------------------------------------------------------
struct PosRef {
    using T = double;
    struct Step {                          // metres per axis per step (parameters::prepare)
        double d = 0.;
        __device__ void set(const WalkArgs &a, uint32_t ts) { d = a.diffusivity[ts]; }
        __device__ bool moves() const { return d != 0.; }          // kernels.cu:130
    };
    __device__ static T start(const WalkArgs &, float x0, int) { return (double)x0; }
    __device__ static float metres(const WalkArgs &, T x, int) { return (float)x; }
    template <bool CROSS>
    __device__ static void step(const WalkArgs &a, float px, float py, float pz, const Step &s, const T xo[3],
                                T xn[3]) {
        move<CROSS>(a, px, py, pz, s.d, xo, xn);
    }
    template <class Rec, bool IDX32>
    __device__ static Vox<IDX32> find(const WalkArgs &a, const T x[3]) { return locate<Rec, IDX32>(a, x); }
    // One attempt (kernels.cu:130-140): the step when the spin moves, then the lookup.
    template <class Rec, bool CROSS, bool IDX32>
    __device__ static Vox<IDX32> advance(const WalkArgs &a, bool moved, float px, float py, float pz,
                                         const Step &s, const T xo[3], T xn[3]) {
        if (moved) step<CROSS>(a, px, py, pz, s, xo, xn);
        return find<Rec, IDX32>(a, xn);
    }
    // kernels.cu:185: the gradient phase of a step, formed in double from float G
    template <class Phase>
    __device__ static void add_grad(Phase &ph, const WalkArgs &a, float Gx, float Gy, float Gz, const T x[3]) {
        ph.add_grad((Gx * x[0] + Gy * x[1] + Gz * x[2]) * 1e-3 * a.timestep_us * 1e-6 * kGamma * kRad2Deg);
    }
};

__device__ __forceinline__ uint32_t cvt_i32_sat_f(float v) {       // truncation, saturating
    uint32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// v < 0 or v >= n, or v is -0.0 (n > 0 finite), as for maybe_outside above.
__device__ __forceinline__ uint32_t maybe_outside_f(float v, float n) {
    return __float_as_uint(v) >= __float_as_uint(n) ? 1u : 0u;
}

// The boundary correction in voxel units for v = po + rw.  A wrap that rounds
// onto n itself (v just below 0) lands on the last float below n.
template <bool CROSS>
__device__ __forceinline__ float fix_axis_vox(float n, float rw, float v) {
    if (CROSS) {
        if (v < 0.f) {
            v += n;
            if (v >= n) v = __uint_as_float(__float_as_uint(n) - 1u);
        } else if (v >= n) {
            v -= n;
        }
        return v;
    }
    const float adj = 2.f * fabsf(rw);
    return v < 0.f ? v + adj : (v >= n ? v - adj : v);
}

struct PosVox {
    using T = float;
    // Voxels per axis per step, times kDrawScale: the production draws come in
    // units of 1 / kDrawScale (draw_bundle).  A tissue that does not move (zero
    // diffusivity, kernels.cu:130) keeps NaN in d[0]: the step formed from it
    // is NaN, which the in-FoV test reads as outside, so such a lane takes the
    // checked path and stays where it is.
    struct Step {
        float d[3] = {__builtin_nanf(""), 0.f, 0.f};
        __device__ void set(const WalkArgs &a, uint32_t ts) {
            const double D = gload<double>(a.diffusivity, (uint64_t)ts * 8) * kDrawScale;
            d[0] = D != 0. ? (float)(D * a.s2g[0]) : __builtin_nanf("");
            d[1] = (float)(D * a.s2g[1]);
            d[2] = (float)(D * a.s2g[2]);
        }
        __device__ bool moves() const { return d[0] == d[0]; }
    };
    __device__ static T start(const WalkArgs &a, float x0, int i) { return x0 * a.s2gf[i]; }
    __device__ static float metres(const WalkArgs &a, T p, int i) { return p * a.vsf[i]; }
    template <class Rec, bool IDX32>
    __device__ static Vox<IDX32> find(const WalkArgs &a, const T p[3]) {
        const uint32_t ix = cvt_i32_sat_f(p[0]), iy = cvt_i32_sat_f(p[1]), iz = cvt_i32_sat_f(p[2]);
        if (ix < (uint32_t)a.nx && iy < (uint32_t)a.ny && iz < (uint32_t)a.nz) {
            Vox<IDX32> v;
            v.ok = true;
            v.rec = (RecIdx<IDX32>)record_of<Rec, IDX32>(a, ix, iy, iz);
            return v;
        }
        return locate_outside<Rec, IDX32>(a, (double)p[0], (double)p[1], (double)p[2]);
    }
    // One attempt: the step when the spin moves, then the lookup.  When every
    // lane of the wave moves to a position whose coordinates all lie in [0, n)
    // (+0 included; a float below the float nearest n truncates below n), the
    // truncations are the subscripts and lie in range, so the record comes
    // straight from the tables.  Otherwise the boundary correction and the
    // checked lookup run, per lane as in step and find.  The test is a ballot,
    // so the branch is uniform across the wave; the step is formed in every
    // lane, and a lane that does not move (a NaN step) keeps po.
    template <class Rec, bool CROSS, bool IDX32>
    __device__ static Vox<IDX32> advance(const WalkArgs &a, bool moved, float px, float py, float pz,
                                         const Step &s, const T po[3], T pn[3]) {
        pn[0] = __builtin_fmaf(px, s.d[0], po[0]);
        pn[1] = __builtin_fmaf(py, s.d[1], po[1]);
        pn[2] = __builtin_fmaf(pz, s.d[2], po[2]);
        const uint32_t out = maybe_outside_f(pn[0], a.nf[0]) | maybe_outside_f(pn[1], a.nf[1]) |
                             maybe_outside_f(pn[2], a.nf[2]);      // a NaN step is outside
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(out != 0) == 0, 1)) {
            Vox<IDX32> v;
            v.ok = true;
            v.rec = (RecIdx<IDX32>)record_of<Rec, IDX32>(a, cvt_i32_sat_f(pn[0]), cvt_i32_sat_f(pn[1]),
                                                         cvt_i32_sat_f(pn[2]));
            return v;
        }
        if (moved) {
            pn[0] = fix_axis_vox<CROSS>(a.nf[0], px * s.d[0], pn[0]);
            pn[1] = fix_axis_vox<CROSS>(a.nf[1], py * s.d[1], pn[1]);
            pn[2] = fix_axis_vox<CROSS>(a.nf[2], pz * s.d[2], pn[2]);
        } else {
            pn[0] = po[0]; pn[1] = po[1]; pn[2] = po[2];           // a spin that does not move stays put
        }
        return find<Rec, IDX32>(a, pn);
    }
    // The gradient phase of a step from the position in metres: in double for the
    // FP64 walk, in float (g2d = 1e-3 dt 1e-6 gamma rad2deg) for the others.
    template <class Phase>
    __device__ static void add_grad(Phase &ph, const WalkArgs &a, float Gx, float Gy, float Gz, const T p[3]) {
        if constexpr (sizeof(typename Phase::R) == 8) {
            const double x = (double)p[0] * a.vs[0], y = (double)p[1] * a.vs[1], z = (double)p[2] * a.vs[2];
            ph.add_grad((Gx * x + Gy * y + Gz * z) * 1e-3 * a.timestep_us * 1e-6 * kGamma * kRad2Deg);
        } else {
            const float x = p[0] * a.vsf[0], y = p[1] * a.vsf[1], z = p[2] * a.vsf[2];
            ph.add((Gx * x + Gy * y + Gz * z) * a.g2d);
        }
    }
};

template <class Rng> using PosOf = typename std::conditional<Rng::kReplay, PosRef, PosVox>::type;

// Per-tissue values the step loop reads when a spin changes tissue: from the
// workgroup's LDS copies (production walks in the 32-bit mode with few
// tissues, tissue_lds), else from the global tables.  LDS layout after the
// per-axis tables: crossing probabilities [n_sub^2], T1 [n_sub], T2 [n_sub] and
// the step sizes [3 n_sub] as PosVox::Step::set forms them.  (Walks in the
// 64-bit mode keep the global tables: measured on the 1000^3 walk, the LDS
// copies made its plain FP32 tiles 12 % slower.)
template <class Rng, bool IDX32>
struct Tissue {
    static constexpr bool kLds = !Rng::kReplay && IDX32;
    const WalkArgs &a;
    __device__ float *tab() const { return reinterpret_cast<float *>(axis_tab) + (uint32_t)(a.nx + a.ny + a.nz); }
    // Copies the tables; every thread of the workgroup calls it, and the caller
    // synchronises the workgroup before the first walk.
    __device__ void fill() const {
        if constexpr (kLds) {
            if (!a.tissue_lds) return;
            float *t = tab();
            const uint32_t n = a.n_sub, nn = n * n;
            for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x) t[i] = a.pXY[i];
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                t[nn + i] = a.T1s[i];
                t[nn + n + i] = a.T2s[i];
                PosVox::Step st;
                st.set(a, i);
                t[nn + 2 * n + 3 * i + 0] = st.d[0];
                t[nn + 2 * n + 3 * i + 1] = st.d[1];
                t[nn + 2 * n + 3 * i + 2] = st.d[2];
            }
        }
    }
    // The global reads end in an empty asm statement, so that the compiler does
    // not merge the two branches into one flat load through a selected pointer.
    __device__ float pxy(uint32_t from, uint32_t to) const {                 // kernels.cu:155
        const uint32_t i = from * a.n_sub + to;
        if (kLds && a.tissue_lds) return tab()[i];
        const float v = gload<float>(a.pXY, (uint64_t)i * 4);
        asm volatile("" ::);
        return v;
    }
    template <class R> __device__ void relax(uint32_t ts, R &T1, R &T2) const {   // kernels.cu:167-168
        if constexpr (sizeof(R) == 8) {
            T1 = (R)a.T1d[ts]; T2 = (R)a.T2d[ts];
        } else {
            if (kLds && a.tissue_lds) {
                const float *t = tab() + a.n_sub * a.n_sub;
                T1 = t[ts]; T2 = t[a.n_sub + ts];
            } else {
                T1 = (R)gload<float>(a.T1s, (uint64_t)ts * 4); T2 = (R)gload<float>(a.T2s, (uint64_t)ts * 4);
                asm volatile("" ::);
            }
        }
    }
    template <class Step> __device__ void step(uint32_t ts, Step &s) const {  // kernels.cu:169
        if constexpr (kLds && std::is_same<Step, PosVox::Step>::value) {
            if (a.tissue_lds) {
                const float *t = tab() + a.n_sub * (a.n_sub + 2) + 3 * ts;
                s.d[0] = t[0]; s.d[1] = t[1]; s.d[2] = t[2];
                return;
            }
        }
        s.set(a, ts);
        asm volatile("" ::);
    }
};

// ---------------------------------------------------------------------------
// One spin.  Returns ST_OK / ST_LOST / ST_EXHAUSTED; counts accepted steps and
// rejected attempts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t next_event(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return min(min(a, b), min(c, d));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    // 64-lane butterfly; carries folded after each level
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t olo = __shfl_xor(lo, o), ohi = __shfl_xor(hi, o);
        const uint32_t nlo = lo + olo;
        hi = hi + ohi + (nlo < lo ? 1u : 0u);
        lo = nlo;
    }
    return ((uint64_t)hi << 32) | lo;
}

template <class Rng, class Rec, class Phase, bool CROSS, bool IDX32>
__device__ int walk_spin(const WalkArgs &a, uint32_t li, uint64_t &n_acc, uint32_t &n_rej) {
    using R = typename Phase::R;
    using P = PosOf<Rng>;
    const uint32_t spin = a.spin_offset + li;
    float *xyz1 = a.XYZ1 + 3 * (uint64_t)li * a.traj_len;
    typename P::T xo[3], xn[3];
    R m[3];
    for (int i = 0; i < 3; i++) {                                  // kernels.cu:95-99
        const float x0 = a.XYZ0[3 * (uint64_t)li + i] * a.xyz_scale;
        xo[i] = xn[i] = P::start(a, x0, i);
        m[i] = (R)a.M0[3 * (uint64_t)li + i];
        if (a.record_traj) xyz1[i] = x0;
    }
    Vox<IDX32> vx = P::template find<Rec, IDX32>(a, xo);           // kernels.cu:102
    if (!vx.ok) {                                                  // the reference reads out of bounds here
        if (!a.record_traj) for (int i = 0; i < 3; i++) xyz1[i] = P::metres(a, xo[i], i);
        return ST_LOST;
    }
    uint32_t ts_old = Rec::label(rec_load<Rec, IDX32>(a.rec, vx.rec));   // :103
    const Tissue<Rng, IDX32> tis{a};
    typename P::Step stp;
    tis.step(ts_old, stp);                                         // :104
    // The reference re-reads T1/T2/diffusivity of ts_old at every voxel change
    // (:167-169); they only change with ts_old, so they are read when it changes.
    // Before the first accepted step T1 = T2 = 0 (:91); no event can use them
    // before that step unless the TR has no time points at all.
    R field = 0, T1 = 0, T2 = 0;
    if (a.n_timepoints > 0) {
        tis.relax(ts_old, T1, T2);
    }
    uint32_t itr = 0;
    Rng rng(a, li, spin);
    rng.peek();
    Phase ph;
    uint64_t slot = 0;
    // A spin that is lost or stuck parks: tp gains the kParked bit, which ends
    // the step loop at its one exit, its natural end, and keeps the count of
    // steps in the low bits; the TR loop then leaves through `done`.  Running
    // out of replayed draws leaves through `done` at once.
    int st = ST_OK;
    uint32_t d = 0, tp = 0;

    for (d = 0; d <= a.n_dummy; ++d) {                              // :107
        const bool last = (d == a.n_dummy);
        {                                                          // :110-120 excitation
            const Excitation e = a.exc[d];
            if (sizeof(R) == 8)
                rot_apply<R>(e.kind_d, (R)a.exc_sd, (R)a.exc_cd, (R)e.spd, (R)e.cpd, m);
            else
                rot_apply<R>(e.kind, (R)a.exc_s, (R)a.exc_c, (R)e.sp, (R)e.cp, m);
        }
        RecIdx<IDX32> rec_old = ~(RecIdx<IDX32>)0;                 // :123-126 (ind_old: no voxel)
        uint32_t old_tp = 0;
        tp = 0;
        uint32_t rf_i = 1, te_i = 0, dp_i = 0, gr_i = 0;
        uint32_t next_rf = (a.n_rf > 1) ? (uint32_t)a.rf[1].t : kNever;
        uint32_t next_te = (last && a.n_te > 0) ? (uint32_t)a.TE_t[0] : kNever;
        uint32_t next_dp = (a.n_deph > 0) ? (uint32_t)a.deph_t[0] : kNever;
        uint32_t next_gr = (a.n_grad > 0) ? (uint32_t)a.grad_t[0] : kNever;
        uint32_t next_ev = next_event(next_rf, next_te, next_dp, next_gr);
        ph.reset();

        while (tp < a.n_timepoints) {                              // :128
            const bool moved = stp.moves();
            if (moved && !rng.avail) { st = ST_EXHAUSTED; goto done; }   // replayed draws used up
            vx = P::template advance<Rec, CROSS, IDX32>(a, moved, rng.px, rng.py, rng.pz, stp, xo, xn);   // :130-140
            bool acc = true;                                       // the attempt becomes a step
            if (!vx.ok) {                                          // :141-147
                tp |= kParked; acc = false;
            } else {
                const bool newvox = vx.rec != rec_old;             // :150
                typename Rec::T r;
                if (newvox)                                        // one gather: label and field
                    r = rec_load<Rec, IDX32>(a.rec, vx.rec);
                // The next attempt's draws while the load flies.  A spin that
                // does not move never uses a draw (its voxel, hence its label,
                // never changes), so the production generator commits in every
                // lane; the replayed stream advances only on a move.
                if (!Rng::kReplay || moved) {
                    rng.commit();
                    rng.peek();
                }
                after_draws(r, rng.px, rng.py, rng.pz);
                if (newvox) {
                    const uint32_t ts = Rec::label(r);             // :152
                    const R fnew = Rec::template field<R>(a, r);   // :166, read speculatively
                    if (ts != ts_old) {                            // :153-163
                        float u;
                        if (!rng.uniform(u)) { st = ST_EXHAUSTED; goto done; }
                        if (u >= tis.pxy(ts_old, ts)) {
                            acc = false;
                            ++n_rej;
                            if (itr++ > a.max_iterations) tp |= kParked;
                        } else {
                            ts_old = ts;
                            tis.relax(ts, T1, T2);
                            tis.step(ts, stp);
                        }
                    }
                    if (acc) {
                        rec_old = vx.rec;                          // :165
                        field = fnew;
                    }
                }
            }
            if (acc) {
                ph.add(field);                                         // :171
                itr = 0;

                if (tp == next_ev) {
                    if (tp == next_dp) {                               // :175-178
                        if (sizeof(R) == 8)
                            ph.add((R)((double)spin * a.deph_deg[dp_i] / (double)a.n_spins_total));
                        else
                            ph.add((R)((float)spin * a.deph_deg[dp_i] / (float)a.n_spins_total));
                        ++dp_i;
                        next_dp = (dp_i < a.n_deph) ? (uint32_t)a.deph_t[dp_i] : kNever;
                    }
                    if (tp == next_gr) {                               // :181-187
                        const float Gx = a.gx[gr_i] * a.gradient_scale;
                        const float Gy = a.gy[gr_i] * a.gradient_scale;
                        const float Gz = a.gz[gr_i] * a.gradient_scale;
                        P::add_grad(ph, a, Gx, Gy, Gz, xn);
                        ++gr_i;
                        next_gr = (gr_i < a.n_grad) ? (uint32_t)a.grad_t[gr_i] : kNever;
                    }
                    if (tp == next_rf) {                               // :190-199
                        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);
                        dephase_relax(m, ph.angle(), T1, T2, te);
                        const RfPulse p = a.rf[rf_i];
                        if (sizeof(R) == 8) rot_apply<R>(p.kind_d, (R)p.sd, (R)p.cd, (R)p.spd, (R)p.cpd, m);
                        else                rot_apply<R>(p.kind, (R)p.s, (R)p.c, (R)p.sp, (R)p.cp, m);
                        ph.reset();
                        old_tp = tp;
                        ++rf_i;
                        next_rf = (rf_i < a.n_rf) ? (uint32_t)a.rf[rf_i].t : kNever;
                    }
                    if (tp == next_te) {                               // :202-215
                        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);
                        dephase_relax(m, ph.angle(), T1, T2, te);
                        const uint64_t o = (uint64_t)li * a.n_te + te_i;
                        R *M1 = (R *)a.M1;
                        M1[3 * o + 0] = m[0]; M1[3 * o + 1] = m[1]; M1[3 * o + 2] = m[2];
                        a.T[o] = (uint8_t)ts_old;
                        ph.reset();
                        old_tp = tp;
                        ++te_i;
                        next_te = (te_i < a.n_te) ? (uint32_t)a.TE_t[te_i] : kNever;
                    }
                    next_ev = next_event(next_rf, next_te, next_dp, next_gr);
                }
                if (a.record_traj) {                                   // :218-221
                    float *dst = xyz1 + 3 * slot;
                    dst[0] = P::metres(a, xn[0], 0); dst[1] = P::metres(a, xn[1], 1); dst[2] = P::metres(a, xn[2], 2);
                    ++slot;
                }
                xo[0] = xn[0]; xo[1] = xn[1]; xo[2] = xn[2];
                ++tp;
            }
        }
        if (tp & kParked) { st = ST_LOST; tp &= ~kParked; goto done; }
        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);   // :226-231
        dephase_relax(m, ph.angle(), T1, T2, te);
    }
    d = a.n_dummy + 1;                                             // every TR done
    tp = 0;
done:
    // Without trajectories XYZ1 holds the last accepted position: the reference
    // stores every accepted position in place (kernels.cu:222-223), and a spin
    // that is lost keeps the one before.  Then the count of accepted steps.
    if (!a.record_traj) for (int i = 0; i < 3; i++) xyz1[i] = P::metres(a, xo[i], i);
    n_acc += (uint64_t)d * a.n_timepoints + tp;
    return st;
}

// One workgroup walks the spins at positions [t0, t0 + spl * blockDim.x) of the
// launch's visiting order (fewer at the end): lane l takes positions
// t0 + k * blockDim.x + l, k = 0 .. spl-1, one after the other.
template <class Rec, class Phase, bool CROSS, bool IDX32, int NV>
__device__ int walk_spin_pairs(const WalkArgs &a, const WalkView *xv, uint32_t li, uint64_t &n_acc, uint32_t &n_rej);

// PAIRS: the production step loop with two attempts per trip (walk_spin_pairs).
template <class Rng, class Rec, class Phase, bool CROSS, bool IDX32, bool PAIRS = false>
__device__ __forceinline__ void walk_block(const WalkArgs &a, uint32_t t0, uint32_t spl = 1) {
    if (IDX32) {
        const uint32_t n_tab = (uint32_t)(a.nx + a.ny + a.nz);
        for (uint32_t i = threadIdx.x; i < n_tab; i += blockDim.x) axis_tab[i] = axis_entry<Rec>(a, i);
        Tissue<Rng, IDX32>{a}.fill();
        __syncthreads();
    }
    uint64_t n_acc = 0;
    uint32_t n_rej = 0, lost = 0, exh = 0;
    for (uint32_t k = 0; k < spl; k++) {
        const uint32_t t = t0 + k * blockDim.x + threadIdx.x;
        if (t < a.n_local) {
            const uint32_t li = a.order ? a.order[t] : t;
            int st;
            if constexpr (Rng::kReplay || !PAIRS) st = walk_spin<Rng, Rec, Phase, CROSS, IDX32>(a, li, n_acc, n_rej);
            else st = walk_spin_pairs<Rec, Phase, CROSS, IDX32, 1>(a, nullptr, li, n_acc, n_rej);
            lost += (st == ST_LOST);
            exh += (st == ST_EXHAUSTED);
        }
    }
    const uint64_t s_acc = wave_sum_u64(n_acc), s_rej = wave_sum_u64(n_rej);
    const uint64_t s_lost = wave_sum_u64(lost), s_exh = wave_sum_u64(exh);
    if ((threadIdx.x & 63) == 0 && a.counters) {
        if (s_lost) atomicAdd(&a.counters[0], (unsigned long long)s_lost);
        if (s_exh)  atomicAdd(&a.counters[1], (unsigned long long)s_exh);
        atomicAdd(&a.counters[2], (unsigned long long)s_acc);
        atomicAdd(&a.counters[3], (unsigned long long)(s_acc + s_rej));
    }
}

// ---------------------------------------------------------------------------
// The fused walk: one trajectory, NV views (FusedArgs).  walk_spin_views is
// walk_spin with the magnetisation, phase, event tables and outputs held per
// view; the single walk keeps its own function so that its code, which sits at
// the register limit, is not perturbed.  Batch tests hold every view to the
// separate walk of its tile, bit for bit.
// ---------------------------------------------------------------------------
// View v of a fused walk: view 0 is the WalkArgs fields, views 1.. are xv[v - 1].
#define SW_VIEW_FIELD(type, name)                                                          \
    __device__ __forceinline__ type view_##name(const WalkArgs &a, const WalkView *xv, int v) { \
        return v == 0 ? a.name : xv[v - 1].name;                                           \
    }
SW_VIEW_FIELD(const RfPulse *, rf)
SW_VIEW_FIELD(const Excitation *, exc)
SW_VIEW_FIELD(const float *, deph_deg)
SW_VIEW_FIELD(const float *, gx)
SW_VIEW_FIELD(const float *, gy)
SW_VIEW_FIELD(const float *, gz)
SW_VIEW_FIELD(float, exc_s)
SW_VIEW_FIELD(float, exc_c)
SW_VIEW_FIELD(double, exc_sd)
SW_VIEW_FIELD(double, exc_cd)
SW_VIEW_FIELD(float, gradient_scale)
SW_VIEW_FIELD(void *, M1)
SW_VIEW_FIELD(float *, XYZ1)
SW_VIEW_FIELD(uint8_t *, T)
SW_VIEW_FIELD(unsigned long long *, counters)
#undef SW_VIEW_FIELD

// One spin.  NV views share the trajectory (draws, moves, lookups, gathers,
// labels and their tissue parameters); each view keeps its own magnetisation
// and phase, formed by the operations of its own walk (walk_spin above).
template <class Rng, class Rec, class Phase, bool CROSS, bool IDX32, int NV>
__device__ int walk_spin_views(const WalkArgs &a, const WalkView *xv, uint32_t li, uint64_t &n_acc,
                               uint32_t &n_rej) {
    using R = typename Phase::R;
    using P = PosOf<Rng>;
    const uint32_t spin = a.spin_offset + li;
    const uint64_t xoff = 3 * (uint64_t)li * a.traj_len;          // this spin's XYZ1 slots
    float *const xyz1_0 = a.XYZ1 + xoff;                           // view 0's, kept in registers
    auto xyz1_of = [&](int v) -> float * { return v == 0 ? xyz1_0 : view_XYZ1(a, xv, v) + xoff; };
    typename P::T xo[3], xn[3];
    R m[NV][3];
    for (int i = 0; i < 3; i++) {                                  // kernels.cu:95-99
        const float x0 = a.XYZ0[3 * (uint64_t)li + i] * a.xyz_scale;
        xo[i] = xn[i] = P::start(a, x0, i);
        #pragma unroll
        for (int v = 0; v < NV; v++) m[v][i] = (R)a.M0[3 * (uint64_t)li + i];
        if (a.record_traj) {
            #pragma unroll
            for (int v = 0; v < NV; v++) xyz1_of(v)[i] = x0;
        }
    }
    Vox<IDX32> vx = P::template find<Rec, IDX32>(a, xo);           // kernels.cu:102
    if (!vx.ok) {                                                  // the reference reads out of bounds here
        if (!a.record_traj) {
            #pragma unroll
            for (int v = 0; v < NV; v++)
                for (int i = 0; i < 3; i++) xyz1_of(v)[i] = P::metres(a, xo[i], i);
        }
        return ST_LOST;
    }
    uint32_t ts_old = Rec::label(rec_load<Rec, IDX32>(a.rec, vx.rec));   // :103
    const Tissue<Rng, IDX32> tis{a};
    typename P::Step stp;
    tis.step(ts_old, stp);                                         // :104
    // The reference re-reads T1/T2/diffusivity of ts_old at every voxel change
    // (:167-169); they only change with ts_old, so they are read when it changes.
    // Before the first accepted step T1 = T2 = 0 (:91); no event can use them
    // before that step unless the TR has no time points at all.
    R field = 0, T1 = 0, T2 = 0;
    if (a.n_timepoints > 0) {
        tis.relax(ts_old, T1, T2);
    }
    uint32_t itr = 0;
    Rng rng(a, li, spin);
    rng.peek();
    Phase ph[NV];
    uint64_t slot = 0;
    // A spin that is lost or stuck parks: tp gains the kParked bit, which ends
    // the step loop at its one exit, its natural end, and keeps the count of
    // steps in the low bits; the TR loop then leaves through `done`.  Running
    // out of replayed draws leaves through `done` at once.
    int st = ST_OK;
    uint32_t d = 0, tp = 0;

    for (d = 0; d <= a.n_dummy; ++d) {                              // :107
        const bool last = (d == a.n_dummy);
        #pragma unroll
        for (int v = 0; v < NV; v++) {                             // :110-120 excitation
            const Excitation e = view_exc(a, xv, v)[d];
            if (sizeof(R) == 8)
                rot_apply<R>(e.kind_d, (R)view_exc_sd(a, xv, v), (R)view_exc_cd(a, xv, v), (R)e.spd, (R)e.cpd, m[v]);
            else
                rot_apply<R>(e.kind, (R)view_exc_s(a, xv, v), (R)view_exc_c(a, xv, v), (R)e.sp, (R)e.cp, m[v]);
        }
        RecIdx<IDX32> rec_old = ~(RecIdx<IDX32>)0;                 // :123-126 (ind_old: no voxel)
        uint32_t old_tp = 0;
        tp = 0;
        uint32_t rf_i = 1, te_i = 0, dp_i = 0, gr_i = 0;
        uint32_t next_rf = (a.n_rf > 1) ? (uint32_t)a.rf[1].t : kNever;
        uint32_t next_te = (last && a.n_te > 0) ? (uint32_t)a.TE_t[0] : kNever;
        uint32_t next_dp = (a.n_deph > 0) ? (uint32_t)a.deph_t[0] : kNever;
        uint32_t next_gr = (a.n_grad > 0) ? (uint32_t)a.grad_t[0] : kNever;
        uint32_t next_ev = next_event(next_rf, next_te, next_dp, next_gr);
        #pragma unroll
        for (int v = 0; v < NV; v++) ph[v].reset();

        while (tp < a.n_timepoints) {                              // :128
            const bool moved = stp.moves();
            if (moved && !rng.avail) { st = ST_EXHAUSTED; goto done; }   // replayed draws used up
            vx = P::template advance<Rec, CROSS, IDX32>(a, moved, rng.px, rng.py, rng.pz, stp, xo, xn);   // :130-140
            bool acc = true;                                       // the attempt becomes a step
            if (!vx.ok) {                                          // :141-147
                tp |= kParked; acc = false;
            } else {
                const bool newvox = vx.rec != rec_old;             // :150
                typename Rec::T r;
                if (newvox)                                        // one gather: label and field
                    r = rec_load<Rec, IDX32>(a.rec, vx.rec);
                // The next attempt's draws while the load flies.  A spin that
                // does not move never uses a draw (its voxel, hence its label,
                // never changes), so the production generator commits in every
                // lane; the replayed stream advances only on a move.
                if (!Rng::kReplay || moved) {
                    rng.commit();
                    rng.peek();
                }
                after_draws(r, rng.px, rng.py, rng.pz);
                if (newvox) {
                    const uint32_t ts = Rec::label(r);             // :152
                    const R fnew = Rec::template field<R>(a, r);   // :166, read speculatively
                    if (ts != ts_old) {                            // :153-163
                        float u;
                        if (!rng.uniform(u)) { st = ST_EXHAUSTED; goto done; }
                        if (u >= tis.pxy(ts_old, ts)) {
                            acc = false;
                            ++n_rej;
                            if (itr++ > a.max_iterations) tp |= kParked;
                        } else {
                            ts_old = ts;
                            tis.relax(ts, T1, T2);
                            tis.step(ts, stp);
                        }
                    }
                    if (acc) {
                        rec_old = vx.rec;                          // :165
                        field = fnew;
                    }
                }
            }
            if (acc) {
                #pragma unroll
                for (int v = 0; v < NV; v++) ph[v].add(field);         // :171
                itr = 0;

                if (tp == next_ev) {
                    if (tp == next_dp) {                               // :175-178
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            const float dd = view_deph_deg(a, xv, v)[dp_i];
                            if (sizeof(R) == 8)
                                ph[v].add((R)((double)spin * dd / (double)a.n_spins_total));
                            else
                                ph[v].add((R)((float)spin * dd / (float)a.n_spins_total));
                        }
                        ++dp_i;
                        next_dp = (dp_i < a.n_deph) ? (uint32_t)a.deph_t[dp_i] : kNever;
                    }
                    if (tp == next_gr) {                               // :181-187
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            const float gs = view_gradient_scale(a, xv, v);
                            const float Gx = view_gx(a, xv, v)[gr_i] * gs;
                            const float Gy = view_gy(a, xv, v)[gr_i] * gs;
                            const float Gz = view_gz(a, xv, v)[gr_i] * gs;
                            P::add_grad(ph[v], a, Gx, Gy, Gz, xn);
                        }
                        ++gr_i;
                        next_gr = (gr_i < a.n_grad) ? (uint32_t)a.grad_t[gr_i] : kNever;
                    }
                    if (tp == next_rf) {                               // :190-199
                        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            dephase_relax(m[v], ph[v].angle(), T1, T2, te);
                            const RfPulse p = view_rf(a, xv, v)[rf_i];
                            if (sizeof(R) == 8) rot_apply<R>(p.kind_d, (R)p.sd, (R)p.cd, (R)p.spd, (R)p.cpd, m[v]);
                            else                rot_apply<R>(p.kind, (R)p.s, (R)p.c, (R)p.sp, (R)p.cp, m[v]);
                            ph[v].reset();
                        }
                        old_tp = tp;
                        ++rf_i;
                        next_rf = (rf_i < a.n_rf) ? (uint32_t)a.rf[rf_i].t : kNever;
                    }
                    if (tp == next_te) {                               // :202-215
                        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);
                        const uint64_t o = (uint64_t)li * a.n_te + te_i;
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            dephase_relax(m[v], ph[v].angle(), T1, T2, te);
                            R *M1 = (R *)view_M1(a, xv, v);
                            M1[3 * o + 0] = m[v][0]; M1[3 * o + 1] = m[v][1]; M1[3 * o + 2] = m[v][2];
                            view_T(a, xv, v)[o] = (uint8_t)ts_old;
                            ph[v].reset();
                        }
                        old_tp = tp;
                        ++te_i;
                        next_te = (te_i < a.n_te) ? (uint32_t)a.TE_t[te_i] : kNever;
                    }
                    next_ev = next_event(next_rf, next_te, next_dp, next_gr);
                }
                if (a.record_traj) {                                   // :218-221
                    #pragma unroll
                    for (int v = 0; v < NV; v++) {
                        float *dst = xyz1_of(v) + 3 * slot;
                        dst[0] = P::metres(a, xn[0], 0); dst[1] = P::metres(a, xn[1], 1);
                        dst[2] = P::metres(a, xn[2], 2);
                    }
                    ++slot;
                }
                xo[0] = xn[0]; xo[1] = xn[1]; xo[2] = xn[2];
                ++tp;
            }
        }
        if (tp & kParked) { st = ST_LOST; tp &= ~kParked; goto done; }
        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);   // :226-231
        #pragma unroll
        for (int v = 0; v < NV; v++) dephase_relax(m[v], ph[v].angle(), T1, T2, te);
    }
    d = a.n_dummy + 1;                                             // every TR done
    tp = 0;
done:
    // Without trajectories XYZ1 holds the last accepted position: the reference
    // stores every accepted position in place (kernels.cu:222-223), and a spin
    // that is lost keeps the one before.  Then the count of accepted steps.
    if (!a.record_traj) {
        #pragma unroll
        for (int v = 0; v < NV; v++)
            for (int i = 0; i < 3; i++) xyz1_of(v)[i] = P::metres(a, xo[i], i);
    }
    n_acc += (uint64_t)d * a.n_timepoints + tp;
    return st;
}

// ---------------------------------------------------------------------------
// The production walk (Philox draws, float voxel-unit positions), NV views:
// walk_spin_views with two attempts per trip of the step loop.  Attempt A moves
// from the current position with its draws; attempt B, the one after it, is
// formed at the same time from A's position with the next draws, and both
// gathers are issued together, so that a wave waits once for two.  B stands when
// A becomes a step, the spin keeps its tissue (hence its step sizes) and the TR
// has a time point left; otherwise B is dropped and made again in the next trip
// from wherever A left the spin, with the same draws.  The draws of attempt k
// are the k-th bundle of the spin whichever trip makes it, so every attempt,
// step, rejection and event happens as in walk_spin_views, bit for bit.
// Both gathers are issued by every lane, a lane that needs none reading record
// 0, so that the loads are straight-line code and the wait for them is one.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void after_draws(T &ra, T &rb, const PhiloxRng::Draw &q2, const PhiloxRng::Draw &q3) {
    after_draws(ra, q2.x, q2.y, q2.z);
    after_draws(rb, q3.x, q3.y, q3.z);
}

template <class Rec, class Phase, bool CROSS, bool IDX32, int NV>
__device__ int walk_spin_pairs(const WalkArgs &a, const WalkView *xv, uint32_t li, uint64_t &n_acc,
                               uint32_t &n_rej) {
    using Rng = PhiloxRng;
    using R = typename Phase::R;
    using P = PosVox;
    using Draw = PhiloxRng::Draw;
    const uint32_t spin = a.spin_offset + li;
    const uint64_t xoff = 3 * (uint64_t)li * a.traj_len;
    float *const xyz1_0 = a.XYZ1 + xoff;
    auto xyz1_of = [&](int v) -> float * { return v == 0 ? xyz1_0 : view_XYZ1(a, xv, v) + xoff; };
    float xo[3], xa[3], xb[3];
    R m[NV][3];
    for (int i = 0; i < 3; i++) {                                  // kernels.cu:95-99
        const float x0 = a.XYZ0[3 * (uint64_t)li + i] * a.xyz_scale;
        xo[i] = xa[i] = xb[i] = P::start(a, x0, i);
        #pragma unroll
        for (int v = 0; v < NV; v++) m[v][i] = (R)a.M0[3 * (uint64_t)li + i];
        if (a.record_traj) {
            #pragma unroll
            for (int v = 0; v < NV; v++) xyz1_of(v)[i] = x0;
        }
    }
    const Vox<IDX32> v0 = P::template find<Rec, IDX32>(a, xo);     // kernels.cu:102
    if (!v0.ok) {
        if (!a.record_traj) {
            #pragma unroll
            for (int v = 0; v < NV; v++)
                for (int i = 0; i < 3; i++) xyz1_of(v)[i] = P::metres(a, xo[i], i);
        }
        return ST_LOST;
    }
    uint32_t ts_old = Rec::label(rec_load<Rec, IDX32>(a.rec, v0.rec));   // :103
    const Tissue<Rng, IDX32> tis{a};
    typename P::Step stp;
    tis.step(ts_old, stp);                                         // :104
    R field = 0, T1 = 0, T2 = 0;
    if (a.n_timepoints > 0) {
        tis.relax(ts_old, T1, T2);
    }
    uint32_t itr = 0;
    const Rng rng(a, li, spin);
    uint64_t att = 0;                                              // attempt of q0
    Draw q0 = rng.draw(0), q1 = rng.draw(1);
    Phase ph[NV];
    uint64_t slot = 0;
    int st = ST_OK;
    uint32_t d = 0, tp = 0;

    for (d = 0; d <= a.n_dummy; ++d) {                              // :107
        const bool last = (d == a.n_dummy);
        #pragma unroll
        for (int v = 0; v < NV; v++) {                             // :110-120 excitation
            const Excitation e = view_exc(a, xv, v)[d];
            if (sizeof(R) == 8)
                rot_apply<R>(e.kind_d, (R)view_exc_sd(a, xv, v), (R)view_exc_cd(a, xv, v), (R)e.spd, (R)e.cpd, m[v]);
            else
                rot_apply<R>(e.kind, (R)view_exc_s(a, xv, v), (R)view_exc_c(a, xv, v), (R)e.sp, (R)e.cp, m[v]);
        }
        RecIdx<IDX32> rec_old = ~(RecIdx<IDX32>)0;                 // :123-126
        uint32_t old_tp = 0;
        tp = 0;
        uint32_t rf_i = 1, te_i = 0, dp_i = 0, gr_i = 0;
        uint32_t next_rf = (a.n_rf > 1) ? (uint32_t)a.rf[1].t : kNever;
        uint32_t next_te = (last && a.n_te > 0) ? (uint32_t)a.TE_t[0] : kNever;
        uint32_t next_dp = (a.n_deph > 0) ? (uint32_t)a.deph_t[0] : kNever;
        uint32_t next_gr = (a.n_grad > 0) ? (uint32_t)a.grad_t[0] : kNever;
        uint32_t next_ev = next_event(next_rf, next_te, next_dp, next_gr);
        #pragma unroll
        for (int v = 0; v < NV; v++) ph[v].reset();

        // One attempt after its gather: kernels.cu:141-221 as in walk_spin_views.
        // Returns whether it became a step; *kept is cleared when the spin
        // changed tissue.
        auto attempt = [&](const Vox<IDX32> &vx, const typename Rec::T &r, bool newvox, const float *xn,
                           uint32_t ubits, bool &kept) -> bool {
            bool acc = true;
            if (!vx.ok) {                                          // :141-147
                tp |= kParked;
                return false;
            }
            if (newvox) {
                const uint32_t ts = Rec::label(r);                 // :152
                const R fnew = Rec::template field<R>(a, r);       // :166
                if (ts != ts_old) {                                // :153-163
                    if (u24_to_unit(ubits) >= tis.pxy(ts_old, ts)) {
                        acc = false;
                        ++n_rej;
                        if (itr++ > a.max_iterations) tp |= kParked;
                    } else {
                        ts_old = ts;
                        tis.relax(ts, T1, T2);
                        tis.step(ts, stp);
                        kept = false;
                    }
                }
                if (acc) {
                    rec_old = vx.rec;                              // :165
                    field = fnew;
                }
            }
            if (acc) {
                #pragma unroll
                for (int v = 0; v < NV; v++) ph[v].add(field);     // :171
                itr = 0;
                if (tp == next_ev) {
                    if (tp == next_dp) {                           // :175-178
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            const float dd = view_deph_deg(a, xv, v)[dp_i];
                            if (sizeof(R) == 8)
                                ph[v].add((R)((double)spin * dd / (double)a.n_spins_total));
                            else
                                ph[v].add((R)((float)spin * dd / (float)a.n_spins_total));
                        }
                        ++dp_i;
                        next_dp = (dp_i < a.n_deph) ? (uint32_t)a.deph_t[dp_i] : kNever;
                    }
                    if (tp == next_gr) {                           // :181-187
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            const float gs = view_gradient_scale(a, xv, v);
                            const float Gx = view_gx(a, xv, v)[gr_i] * gs;
                            const float Gy = view_gy(a, xv, v)[gr_i] * gs;
                            const float Gz = view_gz(a, xv, v)[gr_i] * gs;
                            P::add_grad(ph[v], a, Gx, Gy, Gz, xn);
                        }
                        ++gr_i;
                        next_gr = (gr_i < a.n_grad) ? (uint32_t)a.grad_t[gr_i] : kNever;
                    }
                    if (tp == next_rf) {                           // :190-199
                        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            dephase_relax(m[v], ph[v].angle(), T1, T2, te);
                            const RfPulse p = view_rf(a, xv, v)[rf_i];
                            if (sizeof(R) == 8) rot_apply<R>(p.kind_d, (R)p.sd, (R)p.cd, (R)p.spd, (R)p.cpd, m[v]);
                            else                rot_apply<R>(p.kind, (R)p.s, (R)p.c, (R)p.sp, (R)p.cp, m[v]);
                            ph[v].reset();
                        }
                        old_tp = tp;
                        ++rf_i;
                        next_rf = (rf_i < a.n_rf) ? (uint32_t)a.rf[rf_i].t : kNever;
                    }
                    if (tp == next_te) {                           // :202-215
                        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);
                        const uint64_t o = (uint64_t)li * a.n_te + te_i;
                        #pragma unroll
                        for (int v = 0; v < NV; v++) {
                            dephase_relax(m[v], ph[v].angle(), T1, T2, te);
                            R *M1 = (R *)view_M1(a, xv, v);
                            M1[3 * o + 0] = m[v][0]; M1[3 * o + 1] = m[v][1]; M1[3 * o + 2] = m[v][2];
                            view_T(a, xv, v)[o] = (uint8_t)ts_old;
                            ph[v].reset();
                        }
                        old_tp = tp;
                        ++te_i;
                        next_te = (te_i < a.n_te) ? (uint32_t)a.TE_t[te_i] : kNever;
                    }
                    next_ev = next_event(next_rf, next_te, next_dp, next_gr);
                }
                if (a.record_traj) {                               // :218-221
                    #pragma unroll
                    for (int v = 0; v < NV; v++) {
                        float *dst = xyz1_of(v) + 3 * slot;
                        dst[0] = P::metres(a, xn[0], 0); dst[1] = P::metres(a, xn[1], 1);
                        dst[2] = P::metres(a, xn[2], 2);
                    }
                    ++slot;
                }
                xo[0] = xn[0]; xo[1] = xn[1]; xo[2] = xn[2];
                ++tp;
            }
            return acc;
        };

        while (tp < a.n_timepoints) {                              // :128
            const bool moved = stp.moves();
            const Vox<IDX32> va = P::template advance<Rec, CROSS, IDX32>(a, moved, q0.x, q0.y, q0.z, stp, xo, xa);
            const bool new_a = va.ok && va.rec != rec_old;         // :150
            typename Rec::T ra = rec_load<Rec, IDX32>(a.rec, new_a ? va.rec : (RecIdx<IDX32>)0);
            const Vox<IDX32> vb = P::template advance<Rec, CROSS, IDX32>(a, moved, q1.x, q1.y, q1.z, stp, xa, xb);
            const bool new_b = va.ok && vb.ok && vb.rec != va.rec;
            typename Rec::T rb = rec_load<Rec, IDX32>(a.rec, new_b ? vb.rec : (RecIdx<IDX32>)0);
            const Draw q2 = rng.draw(att + 2), q3 = rng.draw(att + 3);   // while the gathers fly
            after_draws(ra, rb, q2, q3);
            bool kept = true;
            const bool acc_a = attempt(va, ra, new_a, xa, q0.u, kept);
            // After a step A the voxel of the spin is va's, so B's gather is
            // needed exactly when vb differs from it (new_b).
            const bool both = acc_a && kept && tp < a.n_timepoints;
            if (both) attempt(vb, rb, new_b, xb, q1.u, kept);
            att += both ? 2 : 1;
            const Draw n0 = both ? q2 : q1, n1 = both ? q3 : q2;
            q0 = n0; q1 = n1;
        }
        if (tp & kParked) { st = ST_LOST; tp &= ~kParked; goto done; }
        const R te = (R)((tp - old_tp) * (uint32_t)a.timestep_us * 1e-6);   // :226-231
        #pragma unroll
        for (int v = 0; v < NV; v++) dephase_relax(m[v], ph[v].angle(), T1, T2, te);
    }
    d = a.n_dummy + 1;                                             // every TR done
    tp = 0;
done:
    if (!a.record_traj) {
        #pragma unroll
        for (int v = 0; v < NV; v++)
            for (int i = 0; i < 3; i++) xyz1_of(v)[i] = P::metres(a, xo[i], i);
    }
    n_acc += (uint64_t)d * a.n_timepoints + tp;
    return st;
}


// One workgroup walks the spins at positions [t0, t0 + blockDim.x) of the
// launch's visiting order (fewer at the end).
template <class Rng, class Rec, class Phase, bool CROSS, bool IDX32, int NV>
__device__ __forceinline__ void walk_block_views(const WalkArgs &a, const WalkView *xv, uint32_t n_views,
                                                 uint32_t t0, uint32_t spl) {
    if (IDX32) {
        const uint32_t n_tab = (uint32_t)(a.nx + a.ny + a.nz);
        for (uint32_t i = threadIdx.x; i < n_tab; i += blockDim.x) axis_tab[i] = axis_entry<Rec>(a, i);
        Tissue<Rng, IDX32>{a}.fill();
        __syncthreads();
    }
    uint64_t n_acc = 0;
    uint32_t n_rej = 0, lost = 0, exh = 0;
    for (uint32_t k = 0; k < spl; k++) {
        const uint32_t t = t0 + k * blockDim.x + threadIdx.x;
        if (t < a.n_local) {
            const uint32_t li = a.order ? a.order[t] : t;
            const int st = walk_spin_views<Rng, Rec, Phase, CROSS, IDX32, NV>(a, xv, li, n_acc, n_rej);
            lost += (st == ST_LOST);
            exh += (st == ST_EXHAUSTED);
        }
    }
    const uint64_t s_acc = wave_sum_u64(n_acc), s_rej = wave_sum_u64(n_rej);
    const uint64_t s_lost = wave_sum_u64(lost), s_exh = wave_sum_u64(exh);
    if ((threadIdx.x & 63) == 0) {
        for (int v = 0; v < NV; v++) {                 // every view walked these steps
            unsigned long long *c = view_counters(a, xv, v);
            if (c && (v == 0 || (uint32_t)v < n_views)) {
                if (s_lost) atomicAdd(&c[0], (unsigned long long)s_lost);
                if (s_exh)  atomicAdd(&c[1], (unsigned long long)s_exh);
                atomicAdd(&c[2], (unsigned long long)s_acc);
                atomicAdd(&c[3], (unsigned long long)(s_acc + s_rej));
            }
        }
    }
}

// Occupancy target: waves per SIMD the register allocation must allow.
#ifndef SW_WALK_WAVES_PER_EU
#define SW_WALK_WAVES_PER_EU 6
#endif
#ifndef SW_PAIR_WAVES_PER_EU
#define SW_PAIR_WAVES_PER_EU 4
#endif

// The two-attempt loop holds a second position, record and two draw bundles
// more: it is compiled for 4 waves per SIMD (up to 128 registers), which it
// keeps busy with two gathers in flight per wave.
template <class Rng, class Rec, class Phase, bool CROSS, bool IDX32, bool PAIRS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PAIRS ? SW_PAIR_WAVES_PER_EU : SW_WALK_WAVES_PER_EU)))
void walk_kernel(const WalkArgs a) {
    // Blocks are dealt round-robin over the 8 XCDs; remap so that each XCD gets a
    // contiguous range of the visiting order (spatially sorted spins share L2).
    // The grid is a multiple of 8 blocks, so the remap is a permutation.
    const uint32_t per = gridDim.x / 8, b = blockIdx.x;
    const uint32_t lb = (b % 8) * per + b / 8;
    walk_block<Rng, Rec, Phase, CROSS, IDX32, PAIRS>(a, lb * blockDim.x);
}

// The batched walk: block b takes the chunk at slot (b mod 8) * (grid / 8) + b / 8
// of the chunk table, i.e. XCD k works through slots [k * grid/8, (k+1) * grid/8)
// in order.  The host lays each XCD's slots out longest tile first, so the long
// walks start early and the short ones fill in at the end.  NV > 1: fused items
// of up to NV views, whose per-view state costs registers, so the occupancy
// target is one wave lower.
// Occupancy of the fused kernels: the per-view state (magnetisation and one
// phase register per view, two with Kahan compensation) leaves room for 5 waves
// per SIMD, FP32 and Kahan phases alike, with no spill in the step loop; the few
// spills at 5 waves sit in the TR set-up and the trajectory stores.  Measured:
// 64-phantom FP32 batch 69.6 ms at 5, 74.1 ms at 4; Kahan + int16 field 68.5 ms
// at 5, 75.0 ms at 4; 1000^3 bSSFP Kahan + int16 field 486.5 ms at 5, 541.6 ms
// at 4.  (With the exact remainder by 360 written as a loop, the Kahan kernel
// spilled in the step loop at 5 and ran slower there.)
template <class Phase, int NV>
constexpr int walk_waves_per_eu() {
    return NV == 1 ? SW_WALK_WAVES_PER_EU : 5;
}

template <class Rng, class Rec, class Phase, bool CROSS, bool IDX32, int NV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(walk_waves_per_eu<Phase, NV>())))
void walk_batch_kernel(const void *__restrict__ items, const BatchChunk *__restrict__ chunks, uint32_t spl) {
    const uint32_t per = gridDim.x / 8, b = blockIdx.x;
    const BatchChunk c = chunks[(b % 8) * per + b / 8];
    if (c.item == kNoItem) return;                    // padding slot (uniform across the block)
    if constexpr (NV == 1) {
        walk_block<Rng, Rec, Phase, CROSS, IDX32>(((const WalkArgs *)items)[c.item], c.first, spl);
    } else {
        const FusedArgs &f = ((const FusedArgs *)items)[c.item];
        walk_block_views<Rng, Rec, Phase, CROSS, IDX32, NV>(f.a, f.xv, f.n_views, c.first, spl);
    }
}

// 32-bit mode: record byte offsets below 2^32 and the per-axis tables, with the
// tissue tables when the walk keeps them there, within 64 KiB of LDS.
template <class F>
WalkMode mode_of(const WalkArgs &a) {
    uint64_t bytes;
    if constexpr (F::kPacked) {
        bytes = a.super_blocks ? pack_super_count(a.nx, F::pack.bx, kSuper.sx) * a.nb_y * a.nb_z * kSuperLines * 128
                               : pack_brick_count(a.nx, F::pack.bx) * a.nb_y * a.nb_z * 128;
    } else {
        constexpr int lb = F::shape.lbx + F::shape.lby + F::shape.lbz;
        bytes = (brick_count(a.nx, F::shape.lbx) * a.nb_y * a.nb_z << lb) * sizeof(typename F::T);
    }
    const uint64_t n_tab = (uint64_t)a.nx + (uint64_t)a.ny + (uint64_t)a.nz;
    WalkMode m;
    m.idx32 = bytes < (1ull << 32) && (n_tab + tissue_lds_words(a)) * 4 <= 65536;
    m.lds = m.idx32 ? (size_t)(n_tab + tissue_lds_words(a)) * 4 : 0;
    m.bytes = bytes;
    return m;
}

template <class Rng, class F, class P, bool PAIRS>
hipError_t launch_fp_loop(const WalkArgs &a, hipStream_t s, dim3 grid, dim3 block) {
    const WalkMode m = mode_of<F>(a);
    const bool idx32 = m.idx32;
    const size_t lds = m.lds;
    if constexpr (F::kPacked) {
        if (!idx32 && a.super_blocks) return hipErrorInvalidValue;   // super-blocks always fit the 32-bit mode
    }
    if (a.cross_fov) {
        if (idx32) hipLaunchKernelGGL((walk_kernel<Rng, F, P, true, true, PAIRS>), grid, block, lds, s, a);
        else       hipLaunchKernelGGL((walk_kernel<Rng, F, P, true, false, PAIRS>), grid, block, 0, s, a);
    } else {
        if (idx32) hipLaunchKernelGGL((walk_kernel<Rng, F, P, false, true, PAIRS>), grid, block, lds, s, a);
        else       hipLaunchKernelGGL((walk_kernel<Rng, F, P, false, false, PAIRS>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

template <class Rng, class F, class P>
hipError_t launch_fp(const WalkArgs &a, hipStream_t s, dim3 grid, dim3 block) {
    if constexpr (!Rng::kReplay) {
        if (a.pairs) return launch_fp_loop<Rng, F, P, true>(a, s, grid, block);
    }
    return launch_fp_loop<Rng, F, P, false>(a, s, grid, block);
}

template <class Rng, class F, class P, int NV>
hipError_t launch_batch_nv(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                           uint32_t spl, bool cross, int addr, size_t lds, hipStream_t s) {
    const dim3 grid(n_slots), block(threads);
    if constexpr (F::kPacked) {
        if (addr == ADDR_64_MIXED) {
            if (cross) hipLaunchKernelGGL((walk_batch_kernel<Rng, MixedLayout<F>, P, true, false, NV>), grid, block, lds, s, items, chunks, spl);
            else       hipLaunchKernelGGL((walk_batch_kernel<Rng, MixedLayout<F>, P, false, false, NV>), grid, block, lds, s, items, chunks, spl);
            return hipGetLastError();
        }
    }
    const bool idx32 = addr == ADDR_32;
    if (cross) {
        if (idx32) hipLaunchKernelGGL((walk_batch_kernel<Rng, F, P, true, true, NV>), grid, block, lds, s, items, chunks, spl);
        else       hipLaunchKernelGGL((walk_batch_kernel<Rng, F, P, true, false, NV>), grid, block, lds, s, items, chunks, spl);
    } else {
        if (idx32) hipLaunchKernelGGL((walk_batch_kernel<Rng, F, P, false, true, NV>), grid, block, lds, s, items, chunks, spl);
        else       hipLaunchKernelGGL((walk_batch_kernel<Rng, F, P, false, false, NV>), grid, block, lds, s, items, chunks, spl);
    }
    return hipGetLastError();
}

// views: 1, or kMaxViews for fused items (FP32 and Kahan phases only).
template <class Rng, class F, class P>
hipError_t launch_batch_fp(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                           uint32_t spl, int views, bool cross, int addr, size_t lds, hipStream_t s) {
    if (views == 1) return launch_batch_nv<Rng, F, P, 1>(items, chunks, n_slots, threads, spl, cross, addr, lds, s);
    if constexpr (sizeof(typename P::R) == 4) {
        if (views == kMaxViews)
            return launch_batch_nv<Rng, F, P, kMaxViews>(items, chunks, n_slots, threads, spl, cross, addr, lds, s);
    }
    return hipErrorInvalidValue;
}

// The float32 records a walk gathers: the reference layout for replayed draws
// (held bit for bit to the oracle); in production, the packed 5-byte records
// for batches, which are bound by VALU issue and gain from one load a gather,
// and the structure-of-arrays bricks for single walks, which wait on HBM and
// run faster with two aligned loads than with one unaligned one.  Both hold
// the same bricks, so both give the same results bit for bit.
template <class Rng> using RecF32Batch = typename std::conditional<Rng::kReplay, RecF32, RecF32P>::type;
template <class Rng> using RecF32Single = typename std::conditional<Rng::kReplay, RecF32, RecF32S>::type;

template <class Rng>
hipError_t launch_batch_rng(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                            uint32_t spl, int views, int field, int prec, bool cross, int addr, size_t lds,
                            hipStream_t s) {
    if (prec == 2) {
        if (field == 0) return launch_batch_fp<Rng, RecNone, PhaseF64>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 1) return launch_batch_fp<Rng, RecF32Batch<Rng>, PhaseF64>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
    } else if (prec == 1) {
        if (field == 0) return launch_batch_fp<Rng, RecNone, PhaseKahan>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 1) return launch_batch_fp<Rng, RecF32Batch<Rng>, PhaseKahan>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 2) return launch_batch_fp<Rng, RecF16, PhaseKahan>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 3) return launch_batch_fp<Rng, RecBF16, PhaseKahan>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 4) return launch_batch_fp<Rng, RecQ16, PhaseKahan>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
    } else if (prec == 0) {
        if (field == 0) return launch_batch_fp<Rng, RecNone, PhaseF32>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 1) return launch_batch_fp<Rng, RecF32Batch<Rng>, PhaseF32>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 2) return launch_batch_fp<Rng, RecF16, PhaseF32>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 3) return launch_batch_fp<Rng, RecBF16, PhaseF32>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
        if (field == 4) return launch_batch_fp<Rng, RecQ16, PhaseF32>(items, chunks, n_slots, threads, spl, views, cross, addr, lds, s);
    }
    return hipErrorInvalidValue;
}

template <class Rng>
hipError_t launch_rng(const WalkArgs &a, int field, int prec, hipStream_t s) {
    const uint32_t nb = (a.n_local + 255) / 256;
    if (nb == 0) return hipSuccess;
    const dim3 block(256), grid((nb + 7) / 8 * 8);
    if (prec == 2) {
        if (field == 0) return launch_fp<Rng, RecNone, PhaseF64>(a, s, grid, block);
        if (field == 1) return launch_fp<Rng, RecF32Single<Rng>, PhaseF64>(a, s, grid, block);
    } else if (prec == 1) {
        if (field == 0) return launch_fp<Rng, RecNone, PhaseKahan>(a, s, grid, block);
        if (field == 1) return launch_fp<Rng, RecF32Single<Rng>, PhaseKahan>(a, s, grid, block);
        if (field == 2) return launch_fp<Rng, RecF16, PhaseKahan>(a, s, grid, block);
        if (field == 3) return launch_fp<Rng, RecBF16, PhaseKahan>(a, s, grid, block);
        if (field == 4) return launch_fp<Rng, RecQ16, PhaseKahan>(a, s, grid, block);
    } else if (prec == 0) {
        if (field == 0) return launch_fp<Rng, RecNone, PhaseF32>(a, s, grid, block);
        if (field == 1) return launch_fp<Rng, RecF32Single<Rng>, PhaseF32>(a, s, grid, block);
        if (field == 2) return launch_fp<Rng, RecF16, PhaseF32>(a, s, grid, block);
        if (field == 3) return launch_fp<Rng, RecBF16, PhaseF32>(a, s, grid, block);
        if (field == 4) return launch_fp<Rng, RecQ16, PhaseF32>(a, s, grid, block);
    }
    return hipErrorInvalidValue;
}

}  // namespace

// Launchers of the kernel families (one translation unit each).
hipError_t launch_walk_replay(const WalkArgs &a, int field, int prec, hipStream_t s);
hipError_t launch_walk_philox(const WalkArgs &a, int field, int prec, hipStream_t s);
hipError_t launch_walk_batch_replay(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                                    uint32_t spl, int views, int field, int prec, bool cross, int addr, size_t lds,
                                    hipStream_t s);
hipError_t launch_walk_batch_philox(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                                    uint32_t spl, int views, int field, int prec, bool cross, int addr, size_t lds,
                                    hipStream_t s);

}  // namespace sw
