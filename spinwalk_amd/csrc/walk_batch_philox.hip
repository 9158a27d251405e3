// walk_batch_philox.hip -- the production walk kernels (Philox4x32-10 draws)
// of the batched engine: many tiles in one launch, single and fused walks.
#include "walk_impl.hpp"

namespace sw {

hipError_t launch_walk_batch_philox(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                                    uint32_t spl, int views, int field, int prec, bool cross, int addr, size_t lds,
                                    hipStream_t s) {
    return launch_batch_rng<PhiloxRng>(items, chunks, n_slots, threads, spl, views, field, prec, cross, addr, lds, s);
}

}  // namespace sw
