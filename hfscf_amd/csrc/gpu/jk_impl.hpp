// J/K work-item kernel template (one instantiation per angular-momentum class).
#pragma once
#include "kernels.hpp"
#include "quartet.hpp"

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

template <int LA, int LB, int LC, int LD, bool PURE, typename R, typename ACC>
__global__ __launch_bounds__(256) void jk_kernel(JkArgs a) {
    constexpr int NA = PURE && LA > 0 ? 2 * LA + 1 : ncart(LA);
    constexpr int NB = PURE && LB > 0 ? 2 * LB + 1 : ncart(LB);
    constexpr int NC = PURE && LC > 0 ? 2 * LC + 1 : ncart(LC);
    constexpr int ND = PURE && LD > 0 ? 2 * LD + 1 : ncart(LD);
    __shared__ ACC tiles[kRegions * kBlockFuncs * kBlockFuncs];
    __shared__ int rowpre[kMaxRows + 1];
    const WorkItem w = a.items[blockIdx.x];
    const int fA = a.blk_nf[w.A], fB = a.blk_nf[w.B], fC = a.blk_nf[w.C], fD = a.blk_nf[w.D];
    ACC* tJab = tiles;
    ACC* tJcd = tJab + fA * fB;
    ACC* tKac = tJcd + fC * fD;
    ACC* tKad = tKac + fA * fC;
    ACC* tKbc = tKad + fA * fD;
    ACC* tKbd = tKbc + fB * fC;
    const int ntile = fA * fB + fC * fD + fA * fC + fA * fD + fB * fC + fB * fD;
    for (int k = threadIdx.x; k < ntile; k += blockDim.x) tiles[k] = (ACC)0;

    const int nrows = w.r1 - w.r0;
    const int bpo = a.bp_off[w.P], kpo = a.bp_off[w.Q], kn = a.bp_cnt[w.Q];
    const bool same = (w.P == w.Q);
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
        const int i = w.r0 + r;
        int c = ket_count(a.psmean + kpo, kn, a.psmean[bpo + i], a.tol);
        if (same) c = min(c, i + 1);
        rowpre[r + 1] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rowpre[0] = 0;
        for (int r = 0; r < nrows; ++r) rowpre[r + 1] += rowpre[r];
    }
    __syncthreads();
    const int total = rowpre[nrows];
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        int lo = 0, hi = nrows;  // largest r with rowpre[r] <= t
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rowpre[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        const int i = w.r0 + lo;
        const int j = t - rowpre[lo];
        const PairDev bp = a.pairs[bpo + i];
        const PairDev kp = a.pairs[kpo + j];
        double s = 1.0;
        if (bp.sa == bp.sb) s *= 0.5;
        if (kp.sa == kp.sb) s *= 0.5;
        if (same && i == j) s *= 0.5;
        R I[NA * NB * NC * ND];
        eri_block<LA, LB, LC, LD, PURE, R>(bp, kp, a.pp0, a.pp1, a.pp2, a.boys, a.prim_cut, a.psq, a.elem_tol, I);
        digest<NA, NB, NC, ND, R, ACC>(I, s, a.D, a.n, bp, kp, tJab, tJcd, tKac, tKad, tKbc, tKbd, fB, fC, fD);
    }
    __syncthreads();
    double* out = a.slab + w.slab;
    for (int k = threadIdx.x; k < ntile; k += blockDim.x) out[k] = (double)tiles[k];
}

// Launcher of one class (specialised in its own translation unit).
template <int LA, int LB, int LC, int LD>
hipError_t launch_jk_class(bool pure, int mixed, const JkArgs& a, int nitems, hipStream_t s);

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
