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

// Scratch of one dispatch at most kDispatchScratch (the classes that spill run one at a time on
// one stream): fewer persistent workgroups, of 64 lanes when few lanes are allowed.
// A class whose bound is at most kSpillBytes is taken to fit in registers (the (pp|pp) J/K
// kernel, bound 1.6 KB, uses no scratch; the (dd|dd) one, bound 29 KB, spills 13 KB per lane).
constexpr size_t kSpillBytes = 2048;

inline bool class_spills(int la, int lb, int lc, int ld, bool grad) {
    return class_scratch_bound(la, lb, lc, ld, grad) > kSpillBytes;
}

inline LaunchShape scratch_safe_shape(int la, int lb, int lc, int ld, bool grad, LaunchShape want, int nitems) {
    constexpr double kDispatchScratch = 768.0 * 1024.0 * 1024.0;
    const size_t est = class_scratch_bound(la, lb, lc, ld, grad);
    if (est <= kSpillBytes) return want;
    const double lanes = kDispatchScratch / (double)est;
    const int block = want.block > 0 ? want.block : (lanes < 65536.0 ? 64 : 256);
    int grid = want.grid > 0 ? std::min(want.grid, nitems) : nitems;
    grid = std::min(grid, std::max(1, (int)(lanes / block)));
    return LaunchShape{grid, block};
}

}  // namespace hfg
