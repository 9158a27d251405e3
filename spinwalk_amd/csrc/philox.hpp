// philox.hpp -- Philox4x32-10 counter-based generator and the per-attempt draw
// bundle of the production walk.
//
// Philox4x32-10 follows Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3" (SC'11): 10 rounds of two 32x32->64 multiplies by
// 0xD2511F53 / 0xCD9E8D57 with Weyl key bumps 0x9E3779B9 / 0xBB67AE85.
//
// One attempted step of a spin draws one Philox block keyed by the 64-bit seed
// with counter (attempt_lo, attempt_hi, global_spin, 0).  The block gives the
// three displacement normals (two Box-Muller pairs on the top 23 bits of the
// words, the last sine dropped) and, from the low bytes of the first three
// words, the permeability uniform (draw_bundle).  The stream of
// a spin depends only on (seed, spin, attempt), never on the launch shape or on
// how spins are split over launches or devices.  As in the reference, whose
// per-spin generators are re-seeded with seed+spin for every scale
// (src/sim/kernels.cu:77-88), the draws do not depend on the scale index.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SW_HD __host__ __device__ __forceinline__
#else
#define SW_HD inline
#endif

namespace sw {

struct U32x4 { uint32_t v[4]; };

// a ^ b ^ c; one v_bitop3_b32 (truth table 0x96) on the device
SW_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;

SW_HD U32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                          uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = kPhiloxM0, M1 = kPhiloxM1;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
    // The key is uniform across the wave.  Marking it as changed here keeps the
    // compiler from precomputing all 20 round keys outside the caller's loop
    // (20 SGPRs, which spill); the key schedule then runs on the scalar unit.
    asm volatile("" : "+s"(k0), "+s"(k1));
#pragma unroll
#endif
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c0;
        const uint64_t p1 = (uint64_t)M1 * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = xor3(hi1, c1, k0);
        const uint32_t n2 = xor3(hi0, c3, k1);
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
    U32x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// philox4x32_10(c0, c1, c2, c3, k0, k1) with the first round's product
// kPhiloxM1 * c2 given as (m_hi, m_lo): a walk keeps its spin index in c2, so
// that product is formed once per spin instead of once per draw.
SW_HD U32x4 philox4x32_10_pre(uint32_t c0, uint32_t c1, uint32_t m_hi, uint32_t m_lo, uint32_t c3,
                              uint32_t k0, uint32_t k1) {
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(k0), "+s"(k1));
#endif
    {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint32_t n0 = xor3(m_hi, c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c1 = m_lo; c3 = (uint32_t)p0; c0 = n0;
        m_hi = n2;                                  // c2 of round 2
        k0 += W0; k1 += W1;
    }
    uint32_t c2 = m_hi;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int r = 1; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = xor3((uint32_t)(p1 >> 32), c1, k0);
        const uint32_t n2 = xor3((uint32_t)(p0 >> 32), c3, k1);
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += W0; k1 += W1;
    }
    U32x4 o;
    o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// 23-bit uniforms from the top bits of a word, built as floats in [1, 2).
SW_HD float bits_to_float(uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(b);
#else
    float f;
    __builtin_memcpy(&f, &b, 4);
    return f;
#endif
}
SW_HD float u23_open1(uint32_t r)  { return 2.0f - bits_to_float(0x3F800000u | (r >> 9)); }   // (0,1]
SW_HD float u23_closed0(uint32_t r){ return bits_to_float(0x3F800000u | (r >> 9)) - 1.0f; }   // [0,1)
// Permeability uniform from the low bytes of three words: [0,1), 24 bits.
SW_HD float u24_lowbytes(uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t x = (a & 0xFFu) | ((b & 0xFFu) << 8) | ((c & 0xFFu) << 16);
    return (float)x * 5.9604644775390625e-08f;
}

// The 24 bits of u24_lowbytes as an integer (two byte permutes on the device)
// and its value in [0, 1).
SW_HD uint32_t low_bytes3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t ab = __builtin_amdgcn_perm(b, a, 0x0C0C0400u);    // {a.b0, b.b0, 0, 0}
    return __builtin_amdgcn_perm(c, ab, 0x0C040100u);                 // {a.b0, b.b0, c.b0, 0}
#else
    return (a & 0xFFu) | ((b & 0xFFu) << 8) | ((c & 0xFFu) << 16);
#endif
}
SW_HD float u24_to_unit(uint32_t x) { return (float)x * 5.9604644775390625e-08f; }

// 1.0 | r >> 9: the top 23 bits of r as a float in [1, 2) (one v_alignbit).
SW_HD float unit12(uint32_t r) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(__builtin_amdgcn_alignbit(0x7Fu, r, 9));
#else
    return bits_to_float(0x3F800000u | (r >> 9));
#endif
}

// The draws of one attempt from one Philox block r (Box-Muller), in units of
// 1 / sqrt(2 ln 2), the factor the walk folds into its step sizes:
//   x = s1 cos(2 pi f1), y = s1 sin(2 pi f1), z = s3 cos(2 pi f3),
//   s = sqrt(-log2(2 - f)) with f = unit12(r0) (s1) or unit12(r2) (s3),
//   f1 = unit12(r1), f3 = unit12(r3) (the angle in revolutions: cos(2 pi f) of
//   f in [1, 2) is cos of its fraction), u = low_bytes3(r0, r1, r2) / 2^24.
// kDrawScale * s is the Box-Muller radius sqrt(-2 ln(2 - f)), so kDrawScale * x,
// y and z are standard normals.  On the device log2, sqrt, sin and cos are the
// hardware instructions (v_log_f32 is log2; v_sin_f32 and v_cos_f32 take
// revolutions), and the negation rides on the square root's source modifier.
constexpr double kDrawScale = 1.1774100225154747;              // sqrt(2 ln 2)
#if defined(__HIPCC__)
__device__ __forceinline__ void draw_bundle(const U32x4 &r, float &x, float &y, float &z, uint32_t &u) {
    const float l1 = __builtin_amdgcn_logf(2.0f - unit12(r.v[0]));
    const float l2 = __builtin_amdgcn_logf(2.0f - unit12(r.v[2]));
    const float s1 = __builtin_amdgcn_sqrtf(-l1);
    const float s3 = __builtin_amdgcn_sqrtf(-l2);
    const float f1 = unit12(r.v[1]), f3 = unit12(r.v[3]);
    x = s1 * __builtin_amdgcn_cosf(f1);
    y = s1 * __builtin_amdgcn_sinf(f1);
    z = s3 * __builtin_amdgcn_cosf(f3);
    u = low_bytes3(r.v[0], r.v[1], r.v[2]);
}
#endif

}  // namespace sw
