// The fixed device scratch of hfg_build_jk_noalloc (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace hfg {

// 32 MiB per GPU: the first kNoallocCounterDoubles hold the class launches' queue counters, the
// rest the work items' J/K tiles.
constexpr size_t kNoallocScratchDoubles = (size_t)4 << 20;
constexpr size_t kNoallocCounterDoubles = 64;  // 128 int counters

// Device address of this GPU's scratch array (looked up once per device).
hipError_t noalloc_scratch(int device, double** out);

}  // namespace hfg
