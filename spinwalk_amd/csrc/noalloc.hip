// noalloc.hip -- the fixed device scratch of sw_walk_device_noalloc.
//
// sw_walk_device keeps its counters, visiting order and sort buffers in
// engine-owned device memory that grows with hipMalloc on demand.  The
// no-allocation entry point uses this file-scope array instead: every device
// the library runs on holds one copy of it (kNoAllocScratchBytes) from module
// load until the process ends.  Host code never passes `scratch` itself to a
// kernel (on the host that name is only a stand-in for the symbol): the
// device address of each device's copy comes from hipGetSymbolAddress, called
// once per device with that device current, and is kept here.
#include <hip/hip_runtime.h>

#include <mutex>

#include "walk_launch.hpp"

namespace sw {
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void *g_addr[kMaxDevices] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

hipError_t noalloc_scratch_init(int device, void **addr) {
    *addr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;       // the caller makes `device` current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_addr[device] = p;
    }
    *addr = g_addr[device];
    return hipSuccess;
}

}  // namespace sw
