// Scratch-safe launch shapes of the class kernels (used by the run-time dispatchers).
#pragma once
#include <algorithm>

#include "common.hpp"

namespace hfg {

// Per-lane private bytes a class kernel may spill: the used entries of its VRR table (at L+1
// for the gradient) in doubles. The runtime's reported private size cannot be relied on, and a
// dispatch whose resident lanes need more scratch than the device grants fails
// (HSA_STATUS_ERROR_OUT_OF_RESOURCES), so scratch-heavy launches are sized from this bound.
inline size_t class_scratch_bound(int la, int lb, int lc, int ld, bool grad) {
    const int lab = la + lb + (grad ? 1 : 0), lcd = lc + ld + (grad ? 1 : 0), lt = la + lb + lc + ld + (grad ? 1 : 0);
    return sizeof(double) * (size_t)vrr_entries(lab, lcd, lt);
}

// Classes whose kernels run on the device's shared stream, one at a time. Gradient: every class
// whose bound exceeds kSpillBytes (the register kernels of the d classes spill 10-13 KB per
// lane; the workspace kernels share one workspace per context). J/K: the classes whose bound
// exceeds kJkSharedBytes, (dd|dd), (dd|dp) and (dp|dd), which spill 9-14 KB per lane; the
// other J/K classes spill at most 4.3 KB per lane and run concurrently on the side streams.
// (Per-lane private sizes as read from the code-object metadata of the built kernels.)
constexpr size_t kSpillBytes = 2048;
constexpr size_t kJkSharedBytes = 15000;

inline bool class_spills(int la, int lb, int lc, int ld, bool grad) {
    return class_scratch_bound(la, lb, lc, ld, grad) > (grad ? kSpillBytes : kJkSharedBytes);
}

// Gradient launches that spill: at most kDispatchScratch of scratch by the bound, fewer
// persistent workgroups of 64 lanes when few lanes are allowed. J/K launches keep the requested
// shape.
inline LaunchShape scratch_safe_shape(int la, int lb, int lc, int ld, bool grad, LaunchShape want, int nitems) {
    constexpr double kDispatchScratch = 768.0 * 1024.0 * 1024.0;
    const size_t est = class_scratch_bound(la, lb, lc, ld, grad);
    if (!grad || est <= kSpillBytes) return want;
    const double lanes = kDispatchScratch / (double)est;
    const int block = want.block > 0 ? want.block : (lanes < 65536.0 ? 64 : 256);
    int grid = want.grid > 0 ? std::min(want.grid, nitems) : nitems;
    grid = std::min(grid, std::max(1, (int)(lanes / block)));
    return LaunchShape{grid, block};
}

}  // namespace hfg
