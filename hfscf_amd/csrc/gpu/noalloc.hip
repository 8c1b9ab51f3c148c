// Device scratch of hfg_build_jk_noalloc: one file-scope array per GPU that holds the class
// launches' queue counters and the work items' tile slab, so that entry point allocates nothing.
// Its size is fixed; every call on a GPU shares that GPU's copy, which keeps the last call's
// contents (the counters are cleared by a kernel before each use, every slab tile is written
// before it is read), and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "noalloc.hpp"

namespace hfg {

static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

// The device address of this GPU's copy of the array, looked up once per device with that device
// current (the array's name in host code is a host-side stand-in, not a device address).
hipError_t noalloc_scratch(int device, double** out) {
    static std::mutex mu;
    static std::vector<double*> addr;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double*>(p);
    }
    *out = addr[device];
    return hipSuccess;
}

}  // namespace hfg
