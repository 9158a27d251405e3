// probe.hip -- the random-line read probe: the memory ceiling of the voxel gather.
//
// The walk's gather is one 8-byte read from a 128-byte brick line that is,
// for a spin entering a brick, almost never in L2.  This kernel issues only
// such reads: every lane reads 8-byte words at uniformly random 128-byte lines
// of a buffer, with no dependence between reads, so the achieved line rate is
// the ceiling that the memory system sets for a working set of that size
// (L2, Infinity Cache or HBM, as the size decides).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "walk_launch.hpp"

namespace sw {
namespace {

constexpr int kUnroll = 8;

__global__ __launch_bounds__(256) void gather_probe_kernel(const uint64_t *__restrict__ buf, uint64_t n_lines,
                                                           uint32_t n_lanes, uint32_t n_reads, uint64_t seed,
                                                           uint64_t *__restrict__ sink) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_lanes) return;
    // 64-bit LCG per lane (Knuth's MMIX constants); the high bits pick the line
    uint64_t x = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(g + 1));
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n_reads; i += kUnroll) {
        uint64_t v[kUnroll];
#pragma unroll
        for (int k = 0; k < kUnroll; k++) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            const uint64_t line = __umul64hi(x, n_lines);            // in [0, n_lines)
            v[k] = buf[line * 16 + (g & 15)];                        // word g mod 16 of the line
        }
#pragma unroll
        for (int k = 0; k < kUnroll; k++) acc ^= v[k];
    }
    sink[g] = acc;
}

}  // namespace

hipError_t launch_gather_probe(const void *buf, uint64_t bytes, uint32_t n_lanes, uint32_t n_reads, uint64_t seed,
                               uint64_t *sink, hipStream_t s) {
    const uint64_t n_lines = bytes / 128;
    if (n_lines == 0 || n_lanes == 0 || n_reads == 0 || n_reads % kUnroll) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_probe_kernel, dim3((n_lanes + 255) / 256), dim3(256), 0, s,
                       (const uint64_t *)buf, n_lines, n_lanes, n_reads, seed, sink);
    return hipGetLastError();
}

}  // namespace sw
