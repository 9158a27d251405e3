// reduce.hip -- FP64 echo reduction.
//
// The reference reduces on the host in Python: for each scale and echo,
// signal = | mean over spins with T == tissue of (Mx, My) |
// (demo/spinwalk_bold.ipynb:733-736).  Here the sums of Mx, My and the counts
// are formed per (echo, label) in float64 on the device.  Pass 1: block j sums
// a fixed chunk of spins for one (echo, label) pair into partial[pair][j];
// pass 2 adds the partials of each pair in a fixed order (one block per pair).  The order of additions
// is fixed by the launch shape alone, so results are bit-reproducible.
#include <hip/hip_runtime.h>

#include "walk_launch.hpp"

namespace sw {
namespace {

constexpr uint32_t kChunk = 4096;   // spins per block in pass 1
constexpr uint32_t kThreads = 256;

template <class R>
__global__ void reduce_pass1(const R *M1, const uint8_t *T, uint32_t n_spins, uint32_t n_te, uint32_t n_sub,
                             uint32_t n_chunks, double *partial) {
    const uint32_t chunk = blockIdx.x, pair = blockIdx.y;
    const uint32_t te = pair / n_sub, lab = pair % n_sub;
    double sx = 0, sy = 0, sc = 0;
    const uint32_t s0 = chunk * kChunk;
    const uint32_t s1 = min(n_spins, s0 + kChunk);
    for (uint32_t s = s0 + threadIdx.x; s < s1; s += kThreads) {
        const uint64_t o = (uint64_t)s * n_te + te;
        if (T[o] == lab) {
            sx += (double)M1[3 * o + 0];
            sy += (double)M1[3 * o + 1];
            sc += 1.0;
        }
    }
    __shared__ double sh[3][kThreads];
    sh[0][threadIdx.x] = sx; sh[1][threadIdx.x] = sy; sh[2][threadIdx.x] = sc;
    __syncthreads();
    for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *p = partial + 3 * ((uint64_t)pair * n_chunks + chunk);
        p[0] = sh[0][0]; p[1] = sh[1][0]; p[2] = sh[2][0];
    }
}

// One block per (echo, label) pair: thread t adds the partials t, t+256, ... in
// order, then the block's tree adds the 256 thread sums -- the same fixed order
// on every launch.
__global__ void reduce_pass2(const double *partial, uint32_t n_pairs, uint32_t n_chunks, double *out) {
    const uint32_t pair = blockIdx.x;
    double sx = 0, sy = 0, sc = 0;
    for (uint32_t c = threadIdx.x; c < n_chunks; c += kThreads) {
        const double *p = partial + 3 * ((uint64_t)pair * n_chunks + c);
        sx += p[0]; sy += p[1]; sc += p[2];
    }
    __shared__ double sh[3][kThreads];
    sh[0][threadIdx.x] = sx; sh[1][threadIdx.x] = sy; sh[2][threadIdx.x] = sc;
    __syncthreads();
    for (uint32_t w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * pair + 0] = sh[0][0]; out[3 * pair + 1] = sh[1][0]; out[3 * pair + 2] = sh[2][0];
    }
}

}  // namespace

uint64_t reduce_partial_len(uint32_t n_spins, uint32_t n_te, uint32_t n_sub) {
    const uint64_t n_chunks = (n_spins + kChunk - 1) / kChunk;
    return 3ull * n_chunks * n_te * n_sub;
}

hipError_t launch_echo_reduce(const void *M1, int f64, const uint8_t *T, uint32_t n_spins, uint32_t n_te,
                              uint32_t n_sub, double *partial, double *out, hipStream_t s) {
    const uint32_t n_pairs = n_te * n_sub;
    if (n_pairs == 0) return hipSuccess;
    uint32_t n_chunks = (n_spins + kChunk - 1) / kChunk;
    if (n_chunks == 0) {
        // no spins: sums are zero
        hipLaunchKernelGGL(reduce_pass2, dim3(n_pairs), dim3(kThreads), 0, s, partial, n_pairs, 0u, out);
        return hipGetLastError();
    }
    const dim3 g1(n_chunks, n_pairs);
    if (f64)
        hipLaunchKernelGGL((reduce_pass1<double>), g1, dim3(kThreads), 0, s, (const double *)M1, T, n_spins,
                           n_te, n_sub, n_chunks, partial);
    else
        hipLaunchKernelGGL((reduce_pass1<float>), g1, dim3(kThreads), 0, s, (const float *)M1, T, n_spins,
                           n_te, n_sub, n_chunks, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_pass2, dim3(n_pairs), dim3(kThreads), 0, s, partial, n_pairs, n_chunks, out);
    return hipGetLastError();
}

}  // namespace sw
