// walk.hip -- entry points of the walk kernels: dispatch to the kernel families
// (walk_replay.hip, walk_philox.hip, walk_batch_philox.hip), the index mode of a
// phantom, the counter fill kernel and the materialised production draws.
#include "walk_impl.hpp"

namespace sw {

hipError_t launch_walk(const WalkArgs &a, int rng, int field, int prec, hipStream_t s) {
    if (rng == 1) return launch_walk_replay(a, field, prec, s);
    if (rng == 0) return launch_walk_philox(a, field, prec, s);
    return hipErrorInvalidValue;
}

WalkMode walk_mode(const WalkArgs &a, int field, int rng) {
    switch (field) {
    case 0: return mode_of<RecNone>(a);
    case 1: return rng == 1 ? mode_of<RecF32>(a) : mode_of<RecF32P>(a);   // RecF32S: the same bricks
    case 2: return mode_of<RecF16>(a);
    case 4: return mode_of<RecQ16>(a);
    default: return mode_of<RecBF16>(a);
    }
}

hipError_t launch_walk_batch(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                             uint32_t spl, int views, int rng, int field, int prec, bool cross, int addr,
                             size_t lds, hipStream_t s) {
    if (n_slots == 0) return hipSuccess;
    if (n_slots % 8 || threads == 0 || threads > 256 || threads % 64 || spl == 0) return hipErrorInvalidValue;
    if (rng == 1)
        return launch_walk_batch_replay(items, chunks, n_slots, threads, spl, views, field, prec, cross, addr, lds, s);
    if (rng == 0)
        return launch_walk_batch_philox(items, chunks, n_slots, threads, spl, views, field, prec, cross, addr, lds, s);
    return hipErrorInvalidValue;
}

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The production draws, materialised (diagnostics and statistical tests).
// out[(i*n_att + k)*4 + {0,1,2,3}] = {x, y, z, u} of attempt k of spin spin0+i,
// the normals scaled to unit variance.
// ---------------------------------------------------------------------------
__global__ void philox_draws_kernel(uint64_t seed, uint32_t spin0, uint32_t n_spins, uint32_t n_att,
                                    float *out) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (uint64_t)n_spins * n_att) return;
    const uint32_t i = (uint32_t)(g / n_att), k = (uint32_t)(g % n_att);
    WalkArgs a{};
    a.seed = seed;
    PhiloxRng r(a, 0, spin0 + i);
    r.attempt = k;
    r.peek();
    const float c = (float)kDrawScale;                              // the bundle's unit: standard normals
    out[4 * g + 0] = r.px * c; out[4 * g + 1] = r.py * c; out[4 * g + 2] = r.pz * c;
    out[4 * g + 3] = u24_to_unit(r.pu);
}

hipError_t launch_philox_draws(uint64_t seed, uint32_t spin0, uint32_t n_spins, uint32_t n_att,
                               float *out, hipStream_t s) {
    const uint64_t n = (uint64_t)n_spins * n_att;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(philox_draws_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       seed, spin0, n_spins, n_att, out);
    return hipGetLastError();
}

}  // namespace sw
