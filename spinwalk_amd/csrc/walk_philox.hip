// walk_philox.hip -- the production walk kernels (Philox4x32-10 draws), one
// walk per launch.
#include "walk_impl.hpp"

namespace sw {

hipError_t launch_walk_philox(const WalkArgs &a, int field, int prec, hipStream_t s) {
    return launch_rng<PhiloxRng>(a, field, prec, s);
}

}  // namespace sw
