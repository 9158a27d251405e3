// sort.hip -- radix sort of (brick key, spin index) pairs with hipCUB, used to
// build the spatial visiting order of the walk.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "walk_launch.hpp"

namespace sw {

size_t sort_pairs_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                       (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 30);
    return bytes;
}

hipError_t launch_sort_pairs(void *temp, size_t temp_bytes, const uint32_t *kin, uint32_t *kout,
                             const uint32_t *vin, uint32_t *vout, uint32_t n, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, kin, kout, vin, vout, (int)n, 0, 30, s);
}

}  // namespace sw
