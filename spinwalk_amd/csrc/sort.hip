// sort.hip -- the spatial visiting order of the walk.
//
// Spins are visited in the Morton order of the voxel brick (2^lb voxels on a
// side, 8^3 by default) that holds their start position, so that the spins of
// one XCD (a contiguous range of the order) walk in one compact region and
// share its L2.  The order within a brick does
// not matter: a spin's walk does not depend on when it is walked.
//
// When the bricks are few (Morton keys below 2^21) the order is a counting sort:
// a histogram of keys with the rank of each spin in its bin, an exclusive scan,
// and a scatter -- four short kernels.  Larger key ranges take a hipCUB radix
// sort of (key, spin) pairs over the key bits in use.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "walk_launch.hpp"

namespace sw {
namespace {

constexpr int kCountingMaxBits = 21;

__device__ __forceinline__ uint32_t brick_key(const float *XYZ0, float xyz_scale, uint32_t i, double s0, double s1,
                                              double s2, int64_t nx, int64_t ny, int64_t nz, int lb, int bits) {
    const double x = (double)(XYZ0[3 * (uint64_t)i + 0] * xyz_scale) * s0;
    const double y = (double)(XYZ0[3 * (uint64_t)i + 1] * xyz_scale) * s1;
    const double z = (double)(XYZ0[3 * (uint64_t)i + 2] * xyz_scale) * s2;
    const uint32_t mask = (1u << bits) - 1u;
    auto brick = [mask, lb](double v, int64_t n) -> uint32_t {
        // a start position outside the volume sorts with the nearest brick; such
        // a spin is lost at its first lookup anyway
        int64_t b = (v < 9.2e18 && v > -9.2e18) ? (int64_t)v : 0;
        if (b < 0) b = 0;
        if (b >= n) b = n - 1;
        return (uint32_t)(b >> lb) & mask;
    };
    const uint32_t bx = brick(x, nx), by = brick(y, ny), bz = brick(z, nz);
    uint32_t k = 0;
    for (int j = 0; j < bits; j++) {
        k |= ((bz >> j) & 1u) << (3 * j);
        k |= ((by >> j) & 1u) << (3 * j + 1);
        k |= ((bx >> j) & 1u) << (3 * j + 2);
    }
    return k;
}

// Counting sort, pass 1: key of every spin and its rank within its bin.
__global__ void order_count_kernel(const float *XYZ0, float xyz_scale, uint32_t n, double s0, double s1, double s2,
                                   int64_t nx, int64_t ny, int64_t nz, int lb, int bits, uint32_t *key, uint32_t *rank,
                                   uint32_t *hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = brick_key(XYZ0, xyz_scale, i, s0, s1, s2, nx, ny, nz, lb, bits);
    key[i] = k;
    rank[i] = atomicAdd(&hist[k], 1u);
}

// Counting sort, pass 3: every spin to its place.
__global__ void order_scatter_kernel(const uint32_t *key, const uint32_t *rank, const uint32_t *start, uint32_t n,
                                     uint32_t *order) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    order[start[key[i]] + rank[i]] = i;
}

// Radix path: (key, spin index) pairs.
__global__ void order_pairs_kernel(const float *XYZ0, float xyz_scale, uint32_t n, double s0, double s1, double s2,
                                   int64_t nx, int64_t ny, int64_t nz, int lb, int bits, uint32_t *key, uint32_t *val) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = brick_key(XYZ0, xyz_scale, i, s0, s1, s2, nx, ny, nz, lb, bits);
    val[i] = i;
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

OrderPlan order_plan(uint32_t n, int64_t nx, int64_t ny, int64_t nz, int lb) {
    OrderPlan p{};
    p.lb = (uint32_t)lb;
    const int64_t e = (int64_t)1 << lb;
    const int64_t nb = std::max(std::max((nx + e - 1) / e, (ny + e - 1) / e), (nz + e - 1) / e);
    int bits = 1;
    while (bits < 10 && ((int64_t)1 << bits) < nb) bits++;
    p.bits = bits;
    p.counting = 3 * bits <= kCountingMaxBits;
    if (p.counting) {
        p.bins = 1u << (3 * bits);
        size_t scan = 0;
        (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                               (int)p.bins);
        p.temp_bytes = scan;
        p.total_bytes = 2 * align256((size_t)n * 4) + 2 * align256((size_t)p.bins * 4) + align256(scan);
    } else {
        p.bins = 0;
        size_t tmp = 0;
        (void)hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                 (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0,
                                                 3 * bits);
        p.temp_bytes = tmp;
        p.total_bytes = 3 * align256((size_t)n * 4) + align256(tmp);
    }
    return p;
}

hipError_t launch_spatial_order(const OrderPlan &p, void *scratch, const float *XYZ0, float xyz_scale, uint32_t n,
                                const double *s2g, int64_t nx, int64_t ny, int64_t nz, uint32_t *order,
                                hipStream_t s) {
    if (n == 0) return hipSuccess;
    char *base = (char *)scratch;
    const dim3 blk(256), grd((n + 255) / 256);
    if (p.counting) {
        uint32_t *key = (uint32_t *)base;
        uint32_t *rank = (uint32_t *)(base + align256((size_t)n * 4));
        uint32_t *hist = (uint32_t *)(base + 2 * align256((size_t)n * 4));
        uint32_t *start = (uint32_t *)((char *)hist + align256((size_t)p.bins * 4));
        void *temp = (char *)start + align256((size_t)p.bins * 4);
        hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((p.bins + 255) / 256)), blk, 0, s, hist, (uint64_t)p.bins);
        hipLaunchKernelGGL(order_count_kernel, grd, blk, 0, s, XYZ0, xyz_scale, n, s2g[0], s2g[1], s2g[2], nx, ny,
                           nz, (int)p.lb, (int)p.bits, key, rank, hist);
        size_t tb = p.temp_bytes;
        hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, tb, hist, start, (int)p.bins, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(order_scatter_kernel, grd, blk, 0, s, key, rank, start, n, order);
        return hipGetLastError();
    }
    uint32_t *kin = (uint32_t *)base;
    uint32_t *kout = (uint32_t *)(base + align256((size_t)n * 4));
    uint32_t *vin = (uint32_t *)(base + 2 * align256((size_t)n * 4));
    void *temp = base + 3 * align256((size_t)n * 4);
    hipLaunchKernelGGL(order_pairs_kernel, grd, blk, 0, s, XYZ0, xyz_scale, n, s2g[0], s2g[1], s2g[2], nx, ny, nz,
                       (int)p.lb, (int)p.bits, kin, vin);
    size_t tb = p.temp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(temp, tb, kin, kout, vin, order, (int)n, 0, 3 * (int)p.bits, s);
}

}  // namespace sw
