// Pieces shared by the J/K and gradient kernels: pair-list screening helpers, the grid of a
// scratch-bounded persistent launch and the list of angular-momentum classes.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cart.hpp"
#include "kernels.hpp"

namespace hfg {

// Number of ket pairs j (descending mean Schwarz list) with s_i * s_j >= tol.
HFG_INLINE int ket_count(const double* __restrict__ s, int n, double si, double tol) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[mid] * si >= tol)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Quartet t of a work item (flattened over its rows) -> (bra row i, ket column j).
HFG_INLINE void locate(int t, const int* rowpre, int nrows, const WorkItem& w, int& i, int& j) {
    int lo = 0, hi = nrows;  // largest r with rowpre[r] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (rowpre[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    i = w.r0 + lo;
    j = w.c0 + t - rowpre[lo];
}

// Used entries of an Obara-Saika VRR table [e0|f0]^(m), |e| <= LAB, |f| <= LCD, m <= LT - |e| - |f|:
// the largest per-lane array of a class kernel, so a bound on what it can spill to scratch.
HFG_HD int vrr_entries(int LAB, int LCD, int LT) {
    int n = 0;
    for (int a = 0; a <= LAB; ++a)
        for (int b = 0; b <= LCD; ++b)
            if (LT - a - b >= 0) n += ncart(a) * ncart(b) * (LT - a - b + 1);
    return n;
}

// Launch shape of a persistent class launch whose scratch (private segment) stays within
// kScratchBudget: the HIP runtime sizes a dispatch's scratch by the lanes it runs at once, and a
// dispatch beyond the device's scratch limit fails. The private size reported by the runtime
// is used; when it reports none, est_bytes (a per-lane bound from the kernel's arrays) is
// assumed for kernels whose arrays exceed what registers hold. Requested grid 0 = one
// workgroup per work item; a scratch-bound kernel gets 64-lane workgroups so that its few
// resident lanes still spread over every compute unit.
template <typename K>
inline LaunchShape bounded_shape(K kernel, int nitems, LaunchShape want, size_t est_bytes) {
    int block = (want.block > 0 && want.block <= 256) ? want.block : 256;
    int grid = want.grid > 0 ? std::min(want.grid, nitems) : nitems;
    hipFuncAttributes at{};
    size_t priv = 0;
    if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kernel)) == hipSuccess) priv = at.localSizeBytes;
    if (priv == 0 && est_bytes > kRegisterBytes) priv = est_bytes;
    if (priv > 0) {
        const double lanes = kScratchBudget / (double)priv;
        if (lanes < 65536.0 && want.block <= 0) block = 64;
        grid = std::min(grid, std::max(1, (int)(lanes / block)));
    }
    return LaunchShape{std::max(1, grid), block};
}

// Every class with la >= lb, la >= lc >= ld, l <= 2.
#define HFG_FOR_CLASSES(X) \
    X(0, 0, 0, 0)          \
    X(1, 0, 0, 0)          \
    X(1, 0, 1, 0)          \
    X(1, 0, 1, 1)          \
    X(1, 1, 0, 0)          \
    X(1, 1, 1, 0)          \
    X(1, 1, 1, 1)          \
    X(2, 0, 0, 0)          \
    X(2, 0, 1, 0)          \
    X(2, 0, 1, 1)          \
    X(2, 0, 2, 0)          \
    X(2, 0, 2, 1)          \
    X(2, 0, 2, 2)          \
    X(2, 1, 0, 0)          \
    X(2, 1, 1, 0)          \
    X(2, 1, 1, 1)          \
    X(2, 1, 2, 0)          \
    X(2, 1, 2, 1)          \
    X(2, 1, 2, 2)          \
    X(2, 2, 0, 0)          \
    X(2, 2, 1, 0)          \
    X(2, 2, 1, 1)          \
    X(2, 2, 2, 0)          \
    X(2, 2, 2, 1)          \
    X(2, 2, 2, 2)

}  // namespace hfg
