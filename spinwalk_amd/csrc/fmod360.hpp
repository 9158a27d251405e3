// fmod360.hpp -- fmodf(v, 360) without the library's reduction loop.
//
// For |v| < 2^24 the quotient q = trunc(v / 360) is found from one product with
// 1/360 (off by at most one) and fma(-360, q, v) is exact: v and 360 q are both
// multiples of ulp(v) (ulp(v) <= 1), and their difference, smaller than 720 in
// magnitude, is a multiple of its own ulp.  One correction brings the remainder
// into (-360, 360) with the sign of v, which is fmodf's result, and fmodf is
// exact too, so the two agree bit for bit (zero keeps the sign of v).
//
// For |v| >= 2^24, v is an integer m * 2^e (2^23 <= m < 2^24, e >= 1), and
// v mod 360 = (m mod 360) * (2^e mod 360) mod 360 in integers, with
// 2^e mod 360 = 8 * (2^((e - 3) mod 12) mod 45) for e >= 3 (2^12 = 1 mod 45).
// Infinities and NaN give NaN, as fmodf does.  No path loops.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define SW_FMOD_HD __host__ __device__ __forceinline__
#else
#define SW_FMOD_HD inline
#endif

namespace sw {

SW_FMOD_HD float fmod360(float v) {
    const float av = fabsf(v);
    if (!(av < 16777216.f)) {
        if (!(av <= 3.40282347e38f)) return v - v;                 // NaN or infinity
        uint32_t u;
        memcpy(&u, &av, 4);
        const uint32_t m = (u & 0x7FFFFFu) | 0x800000u;
        const uint32_t e = (u >> 23) - 150u;                       // av = m * 2^e, e >= 1
        const uint32_t p2 = e >= 3u ? 8u * ((1u << ((e - 3u) % 12u)) % 45u) : (1u << e);
        return copysignf((float)((m % 360u) * p2 % 360u), v);
    }
    const float q = truncf(v * (1.f / 360.f));
    float r = fmaf(-360.f, q, v);
    if (v >= 0.f) {
        if (r < 0.f) r += 360.f;
        else if (r >= 360.f) r -= 360.f;
    } else {
        if (r > 0.f) r -= 360.f;
        else if (r <= -360.f) r += 360.f;
    }
    return r == 0.f ? copysignf(0.f, v) : r;
}

}  // namespace sw
