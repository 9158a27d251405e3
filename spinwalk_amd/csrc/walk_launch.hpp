// walk_launch.hpp -- host-callable launchers of the HIP kernels (defined in the
// .hip files, called from engine.cpp).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "walk_args.hpp"

namespace sw {

// The walk.  rng: 0 Philox, 1 replay.  field: 0 none, 1 f32, 2 f16, 3 bf16, 4 q16.
// prec: 0 f32, 1 f32+Kahan, 2 f64.
hipError_t launch_walk(const WalkArgs &a, int rng, int field, int prec, hipStream_t s);

// The index mode the walk takes for a phantom and field format: 32-bit record
// offsets with per-axis LDS tables of `lds` bytes, or 64-bit arithmetic.
struct WalkMode { bool idx32; size_t lds; uint64_t bytes; };   // bytes: the gathered records
WalkMode walk_mode(const WalkArgs &a, int field, int rng);

// The batched walk over a chunk table of n_slots entries (a multiple of 8),
// one workgroup of `threads` lanes (64, 128 or 256) per chunk of threads * spl
// spins (each lane walks spl spins, one after the other):
// items[] (WalkArgs, or FusedArgs when views == kMaxViews; prec 0 or 1 only)
// and chunks[] in device memory; every item shares rng, field, prec, cross and
// the address mode (AddrMode; ADDR_32 with `lds` bytes covering every item's
// tables).
hipError_t launch_walk_batch(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                             uint32_t spl, int views, int rng, int field, int prec, bool cross, int addr,
                             size_t lds, hipStream_t s);

// Production draws of spins [spin0, spin0+n_spins), attempts [0, n_att).
hipError_t launch_philox_draws(uint64_t seed, uint32_t spin0, uint32_t n_spins, uint32_t n_att,
                               float *out, hipStream_t s);

// Random-line read probe (probe.hip): n_lanes lanes, n_reads (a multiple of 8)
// 8-byte reads each at uniformly random 128-byte lines of buf[0, bytes);
// sink[n_lanes] keeps the reads live.
hipError_t launch_gather_probe(const void *buf, uint64_t bytes, uint32_t n_lanes, uint32_t n_reads, uint64_t seed,
                               uint64_t *sink, hipStream_t s);

// Phantom preparation: max label and max |field| (into stats[0], stats[1] as
// uint32 bit patterns; stats must be zeroed by the caller).
hipError_t launch_phantom_stats(const uint8_t *mask, const float *field, uint64_t n,
                                uint32_t *stats, hipStream_t s);
// Row-major mask/field -> bricked voxel records: kind 0 Rec8 f32, 3 Rec4 label
// only; kinds 1 fp16, 2 bf16 and 4 int16 fixed point as packed 3-byte records
// (kPack16).  inv_scale divides the field before the 16-bit kinds round it.
// `out` must be zero-filled for padding.
hipError_t launch_records(const uint8_t *mask, const float *field, uint64_t nx, uint64_t ny, uint64_t nz, int kind,
                          float inv_scale, void *out, hipStream_t s);
// Row-major mask/field -> the float32 bricks of production walks: the packed
// 5-byte records (kPackF32) batches gather, or with soa the structure-of-arrays
// records (kSoaF32) single walks gather; `out` must be zero-filled for padding.
hipError_t launch_f32_records(const uint8_t *mask, const float *field, uint64_t nx, uint64_t ny, uint64_t nz,
                              bool soa, void *out, hipStream_t s);
// FP64 echo reduction into out[(te*n_sub+label)*3 + {0,1,2}]; partial is scratch
// of reduce_partial_len(n_spins, n_te, n_sub) doubles.
uint64_t reduce_partial_len(uint32_t n_spins, uint32_t n_te, uint32_t n_sub);
hipError_t launch_echo_reduce(const void *M1, int f64, const uint8_t *T, uint32_t n_spins,
                              uint32_t n_te, uint32_t n_sub, double *partial, double *out,
                              hipStream_t s);

// The fixed scratch of the no-allocation walk (noalloc.hip): its size, and the
// device address of `device`'s copy, looked up once per device (`device` must
// be current on the first call for it).
size_t noalloc_scratch_bytes();
hipError_t noalloc_scratch_init(int device, void **addr);

// Clears n counters with a fill kernel (no memset node, safe under graph capture).
hipError_t launch_zero_u64(unsigned long long *p, int n, hipStream_t s);

// The spatial visiting order (sort.hip): order[k] = the spin walked k-th, sorted
// by the Morton index of the (2^lb)^3 voxel brick of its start position (lb 3
// by default).  order_plan sizes the scratch (total_bytes) and picks a counting
// sort or a radix sort.
constexpr int kOrderBrickLog2 = 3;
struct OrderPlan {
    uint32_t lb;            // brick edge 2^lb voxels
    bool counting;
    uint32_t bits;          // Morton bits per axis
    uint32_t bins;          // counting sort: 2^(3 bits)
    size_t temp_bytes;      // hipCUB scan or sort temp
    size_t total_bytes;     // scratch needed
};
OrderPlan order_plan(uint32_t n, int64_t nx, int64_t ny, int64_t nz, int lb = kOrderBrickLog2);
hipError_t launch_spatial_order(const OrderPlan &p, void *scratch, const float *XYZ0, float xyz_scale, uint32_t n,
                                const double *s2g, int64_t nx, int64_t ny, int64_t nz, uint32_t *order,
                                hipStream_t s);

}  // namespace sw
