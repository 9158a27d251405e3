// walk_replay.hip -- the walk kernels fed the reference CPU path's recorded
// draws (single walks and batches): the replay that is held bit for bit to the
// oracle.
#include "walk_impl.hpp"

namespace sw {

hipError_t launch_walk_replay(const WalkArgs &a, int field, int prec, hipStream_t s) {
    return launch_rng<ReplayRng>(a, field, prec, s);
}

hipError_t launch_walk_batch_replay(const void *items, const BatchChunk *chunks, uint32_t n_slots, uint32_t threads,
                                    uint32_t spl, int views, int field, int prec, bool cross, int addr, size_t lds,
                                    hipStream_t s) {
    return launch_batch_rng<ReplayRng>(items, chunks, n_slots, threads, spl, views, field, prec, cross, addr, lds, s);
}

}  // namespace sw
