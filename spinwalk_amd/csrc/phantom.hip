// phantom.hip -- phantom preparation kernels: label count, field range and the
// 16-bit copies of the fieldmap.
//
// The reference keeps the fieldmap in float32 (src/sim/monte_carlo.cu:98-123,
// 254).  The mixed-precision walk gathers a 16-bit copy:
// - float16 or bfloat16 of field_T / half_scale, where half_scale is a power of
//   two chosen so that the largest |field| lands near 2^14: scaling by a power of
//   two is exact, and it keeps the 1/r^2 and 1/r^3 tails of vessel fields out of
//   the float16 subnormal range for as long as possible.  The error is relative:
//   up to 2^-11 of the value for float16;
// - int16 fixed point rint(field_T / q_scale), q_scale = max|field| / 32767.
//   The error is absolute, at most max|field| / 65534, so every value above 3 %
//   of the largest is held more closely than in float16; a field that is one
//   constant inside every inclusion (cylinders along B0) is held exactly at its
//   largest value.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include "walk_launch.hpp"

namespace sw {
namespace {

__global__ void stats_kernel(const uint8_t *mask, const float *field, uint64_t n, uint32_t *stats) {
    uint32_t mlab = 0, mabs = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        mlab = max(mlab, (uint32_t)mask[i]);
        if (field) mabs = max(mabs, __float_as_uint(fabsf(field[i])));   // |f| bits order like values
    }
    for (int o = 32; o > 0; o >>= 1) {
        mlab = max(mlab, (uint32_t)__shfl_xor((int)mlab, o));
        mabs = max(mabs, (uint32_t)__shfl_xor((int)mabs, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&stats[0], mlab);
        atomicMax(&stats[1], mabs);
    }
}

// The 16-bit field of a voxel: KIND 1 fp16, 2 bf16 of field_T / half_scale;
// 4 int16 fixed point rint(field_T / q_scale).
template <int KIND>
__device__ uint32_t field16(const float *field, uint64_t v, float inv_scale) {
    uint32_t h = 0;
    if (KIND == 1 && field) h = __half_as_ushort(__float2half_rn(field[v] * inv_scale));
    if (KIND == 2 && field) {
        const uint32_t bits = __float_as_uint(field[v] * inv_scale);
        h = (bits + 0x7FFFu + ((bits >> 16) & 1u)) >> 16;          // round to nearest even
    }
    if (KIND == 4 && field) {
        // nearest step, ties to even; |q| <= 32767 up to the rounding of inv_scale
        const double q = rint((double)field[v] * (double)inv_scale);
        h = (uint32_t)(int32_t)fmin(32767.0, fmax(-32767.0, q)) & 0xFFFFu;
    }
    return h & 0xFFFFu;
}

// Row-major mask/field -> bricked records (see walk_args.hpp).  KIND 0: Rec8
// {float field_T, label}; 3: Rec4 label only.
template <int LX, int LY, int LZ, int KIND>
__global__ void records_kernel(const uint8_t *mask, const float *field, uint64_t n, uint32_t ny, uint32_t nz,
                               uint64_t nb_y, uint64_t nb_z, float inv_scale, void *out) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t iz = v % nz, iy = (v / nz) % ny, ix = v / ((uint64_t)nz * ny);
        const uint64_t b = ((ix >> LX) * nb_y + (iy >> LY)) * nb_z + (iz >> LZ);
        const uint64_t r = (b << (LX + LY + LZ)) | ((ix & ((1u << LX) - 1)) << (LY + LZ)) |
                           ((iy & ((1u << LY) - 1)) << LZ) | (iz & ((1u << LZ) - 1));
        const uint32_t lab = mask[v];
        if (KIND == 0) ((uint2 *)out)[r] = make_uint2(__float_as_uint(field ? field[v] : 0.f), lab);
        else ((uint32_t *)out)[r] = lab << 16;
    }
}

// Row-major mask/field -> packed 3-byte 16-bit records (kPack16).
template <int KIND>
__global__ void packed16_records_kernel(const uint8_t *mask, const float *field, uint64_t n, uint32_t ny,
                                        uint32_t nz, uint64_t nb_y, uint64_t nb_z, bool sup, float inv_scale,
                                        uint8_t *out) {
    constexpr PackShape S = kPack16;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t iz = v % nz, iy = (v / nz) % ny, ix = v / ((uint64_t)nz * ny);
        const uint64_t b = pack_line(ix / S.bx, iy / S.by, iz / S.bz, nb_y, nb_z, sup);
        const uint32_t j = (uint32_t)(((ix % S.bx) * S.by + iy % S.by) * S.bz + iz % S.bz);
        uint8_t *r = out + b * 128 + S.rec_bytes * j;
        const uint32_t h = field16<KIND>(field, v, inv_scale);
        r[0] = mask[v];
        r[1] = (uint8_t)h;
        r[2] = (uint8_t)(h >> 8);
    }
}

// Row-major mask/field -> float32 bricks of 2 x 3 x 4 voxels: packed 5-byte
// records (kPackF32) or structure-of-arrays records (kSoaF32).
template <bool SOA>
__global__ void f32_records_kernel(const uint8_t *mask, const float *field, uint64_t n, uint32_t ny, uint32_t nz,
                                   uint64_t nb_y, uint64_t nb_z, bool sup, uint8_t *out) {
    constexpr PackShape S = SOA ? kSoaF32 : kPackF32;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t iz = v % nz, iy = (v / nz) % ny, ix = v / ((uint64_t)nz * ny);
        const uint64_t b = pack_line(ix / S.bx, iy / S.by, iz / S.bz, nb_y, nb_z, sup);
        const uint32_t j = (uint32_t)(((ix % S.bx) * S.by + iy % S.by) * S.bz + iz % S.bz);
        const uint32_t f = __float_as_uint(field ? field[v] : 0.f);
        if constexpr (SOA) {
            uint8_t *brick = out + b * 128;
            brick[j] = mask[v];
            *(uint32_t *)(brick + S.nvox + 4 * j) = f;
        } else {
            uint8_t *r = out + b * 128 + S.rec_bytes * j;
            r[0] = mask[v];
            r[1] = (uint8_t)f; r[2] = (uint8_t)(f >> 8); r[3] = (uint8_t)(f >> 16); r[4] = (uint8_t)(f >> 24);
        }
    }
}

}  // namespace

hipError_t launch_f32_records(const uint8_t *mask, const float *field, uint64_t nx, uint64_t ny, uint64_t nz,
                              bool soa, void *out, hipStream_t s) {
    const uint64_t n = nx * ny * nz;
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    const bool sup = pack_super(nx, ny, nz, kPackF32);
    const uint64_t nby = pack_pitch_y(ny, kPackF32, sup), nbz = pack_pitch_z(nz, kPackF32, sup);
    if (soa)
        hipLaunchKernelGGL(f32_records_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, s, mask, field, n,
                           (uint32_t)ny, (uint32_t)nz, nby, nbz, sup, (uint8_t *)out);
    else
        hipLaunchKernelGGL(f32_records_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, s, mask, field, n,
                           (uint32_t)ny, (uint32_t)nz, nby, nbz, sup, (uint8_t *)out);
    return hipGetLastError();
}

hipError_t launch_phantom_stats(const uint8_t *mask, const float *field, uint64_t n, uint32_t *stats,
                                hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(stats_kernel, dim3((unsigned)blocks), dim3(256), 0, s, mask, field, n, stats);
    return hipGetLastError();
}

hipError_t launch_records(const uint8_t *mask, const float *field, uint64_t nx, uint64_t ny, uint64_t nz, int kind,
                          float inv_scale, void *out, hipStream_t s) {
    const uint64_t n = nx * ny * nz;
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    const dim3 g((unsigned)blocks), b(256);
    if (kind == 0) {
        const uint64_t nby = brick_count(ny, kBrick8.lby), nbz = brick_count(nz, kBrick8.lbz);
        hipLaunchKernelGGL((records_kernel<kBrick8.lbx, kBrick8.lby, kBrick8.lbz, 0>), g, b, 0, s, mask, field, n,
                           (uint32_t)ny, (uint32_t)nz, nby, nbz, inv_scale, out);
    } else if (kind == 3) {
        const uint64_t nby = brick_count(ny, kBrick4.lby), nbz = brick_count(nz, kBrick4.lbz);
        hipLaunchKernelGGL((records_kernel<kBrick4.lbx, kBrick4.lby, kBrick4.lbz, 3>), g, b, 0, s, mask, field, n,
                           (uint32_t)ny, (uint32_t)nz, nby, nbz, inv_scale, out);
    } else {
        const bool sup = pack_super(nx, ny, nz, kPack16);
        const uint64_t nby = pack_pitch_y(ny, kPack16, sup), nbz = pack_pitch_z(nz, kPack16, sup);
        uint8_t *o = (uint8_t *)out;
        if (kind == 1)
            hipLaunchKernelGGL((packed16_records_kernel<1>), g, b, 0, s, mask, field, n, (uint32_t)ny, (uint32_t)nz,
                               nby, nbz, sup, inv_scale, o);
        else if (kind == 2)
            hipLaunchKernelGGL((packed16_records_kernel<2>), g, b, 0, s, mask, field, n, (uint32_t)ny, (uint32_t)nz,
                               nby, nbz, sup, inv_scale, o);
        else
            hipLaunchKernelGGL((packed16_records_kernel<4>), g, b, 0, s, mask, field, n, (uint32_t)ny, (uint32_t)nz,
                               nby, nbz, sup, inv_scale, o);
    }
    return hipGetLastError();
}

}  // namespace sw
