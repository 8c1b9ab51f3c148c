// J/K work-item kernel template (one instantiation per angular-momentum class).
#pragma once
#include "kernels.hpp"
#include <algorithm>

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

// Density-weighted Schwarz bound of a quartet, the form of hfscf's (disabled) direct digest
// (hfscf: src/scfsolver/hfscf_hf_solver.cpp:317-324) on shell blocks.
HFG_INLINE double density_bound(const PairDev& bp, const PairDev& kp, const float* __restrict__ dm, int ns) {
    const int a = bp.sa, b = bp.sb, c = kp.sa, d = kp.sb;
    float x = 4.0f * fmaxf(dm[a * ns + b], dm[c * ns + d]);
    x = fmaxf(x, fmaxf(dm[a * ns + c], dm[a * ns + d]));
    x = fmaxf(x, fmaxf(dm[b * ns + c], dm[b * ns + d]));
    return bp.sq_max * kp.sq_max * (double)x;
}

constexpr int kChunk = 1024;  // candidate quartets screened per pass when density screening is on

template <int LA, int LB, int LC, int LD, bool PURE, typename R, typename ACC>
__global__ __launch_bounds__(256) void jk_kernel(JkArgs a) {
    constexpr int NA = PURE && LA > 0 ? 2 * LA + 1 : ncart(LA);
    constexpr int NB = PURE && LB > 0 ? 2 * LB + 1 : ncart(LB);
    constexpr int NC = PURE && LC > 0 ? 2 * LC + 1 : ncart(LC);
    constexpr int ND = PURE && LD > 0 ? 2 * LD + 1 : ncart(LD);
    __shared__ ACC tiles[kRegions * kBlockFuncs * kBlockFuncs];
    __shared__ int rowpre[kMaxRows + 1];
    __shared__ int clist[kChunk];
    __shared__ int ccount;
    const bool dscreen = a.dtol > 0.0;
    for (int item = blockIdx.x; item < a.nitems; item += gridDim.x) {
        const WorkItem w = a.items[item];
        const int fA = a.blk_nf[w.A], fB = a.blk_nf[w.B], fC = a.blk_nf[w.C], fD = a.blk_nf[w.D];
        ACC* tJab = tiles;
        ACC* tJcd = tJab + fA * fB;
        ACC* tKac = tJcd + fC * fD;
        ACC* tKad = tKac + fA * fC;
        ACC* tKbc = tKad + fA * fD;
        ACC* tKbd = tKbc + fB * fC;
        const int ntile = fA * fB + fC * fD + fA * fC + fA * fD + fB * fC + fB * fD;
        __syncthreads();  // the previous item's flush has read the tiles
        for (int k = threadIdx.x; k < ntile; k += blockDim.x) tiles[k] = (ACC)0;
        const int nrows = w.r1 - w.r0;
        const int bpo = a.bp_off[w.P], kpo = a.bp_off[w.Q], kn = a.bp_cnt[w.Q];
        const bool same = (w.P == w.Q);
        for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
            const int i = w.r0 + r;
            int c = ket_count(a.psmean + kpo, kn, a.psmean[bpo + i], a.tol);
            if (same) c = min(c, i + 1);
            rowpre[r + 1] = max(0, min(c, w.c1) - w.c0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            rowpre[0] = 0;
            for (int r = 0; r < nrows; ++r) rowpre[r + 1] += rowpre[r];
        }
        __syncthreads();
        const int total = rowpre[nrows];
        // total and dscreen are uniform over the workgroup, so every thread takes the same
        // number of passes and reaches the same barriers
        for (int base = 0; base < total; base += kChunk) {
            const int cnt = min(kChunk, total - base);
            int ntodo = cnt;
            if (dscreen) {
                if (threadIdx.x == 0) ccount = 0;
                __syncthreads();
                for (int u = threadIdx.x; u < cnt; u += blockDim.x) {
                    int i, j;
                    locate(base + u, rowpre, nrows, w, i, j);
                    if (density_bound(a.pairs[bpo + i], a.pairs[kpo + j], a.dmax, a.nshell) >= a.dtol)
                        clist[atomicAdd(&ccount, 1)] = base + u;
                }
                __syncthreads();
                ntodo = ccount;
            }
            for (int u = threadIdx.x; u < ntodo; u += blockDim.x) {
                int i, j;
                locate(dscreen ? clist[u] : base + u, rowpre, nrows, w, i, j);
                const PairDev bp = a.pairs[bpo + i];
                const PairDev kp = a.pairs[kpo + j];
                double s = 1.0;
                if (bp.sa == bp.sb) s *= 0.5;
                if (kp.sa == kp.sb) s *= 0.5;
                if (same && i == j) s *= 0.5;
                R I[NA * NB * NC * ND];
                eri_block<LA, LB, LC, LD, PURE, R>(bp, kp, a.pp0, a.pp1, a.pp2, a.boys, a.prim_cut, a.psq, a.elem_tol,
                                                   I);
                digest<NA, NB, NC, ND, R, ACC>(I, s, a.D, a.n, bp, kp, tJab, tJcd, tKac, tKad, tKbc, tKbd, fB, fC,
                                               fD);
            }
            if (dscreen) __syncthreads();  // every thread has read ccount and clist
        }
        __syncthreads();
        double* out = a.slab + w.slab;
        for (int k = threadIdx.x; k < ntile; k += blockDim.x) out[k] = (double)tiles[k];
    }
}

// Launchers of one class, FP64 and mixed (each specialised in its own translation unit).
template <int LA, int LB, int LC, int LD>
hipError_t launch_jk_class64(bool pure, const JkArgs& a, int nitems, hipStream_t s);
template <int LA, int LB, int LC, int LD>
hipError_t launch_jk_class32(bool pure, const JkArgs& a, int nitems, hipStream_t s);

// Workgroups for a launch of `kernel` over nitems work items: one per item, unless the
// kernel's scratch would then exceed kScratchBudget for one dispatch.
template <typename K>
inline int grid_for(K kernel, int nitems, int block) {
    hipFuncAttributes at{};
    if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(kernel)) != hipSuccess) return nitems;
    const double per_wg = (double)at.localSizeBytes * block;
    if (per_wg <= 0.0) return nitems;
    const int cap = (int)(kScratchBudget / per_wg);
    return std::max(1, std::min(nitems, cap));
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
