// J/K work-item kernel template (one instantiation per angular-momentum class).
#pragma once
#include "common.hpp"
#include "quartet.hpp"

namespace hfg {

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

// Writes every canonical element (i >= j, k >= l, ij >= kl; hfscf's index_ijkl, hfscf:
// src/math/hfscf_math.hpp:90) of one quartet's output-basis block I to the packed vector, once:
// each unordered pair of shell pairs is one quartet of the lists, so only a shell paired with
// itself (i < j within it) or a shell pair paired with itself (ij < kl) repeats an element.
template <int NA, int NB, int NC, int ND, typename R>
HFG_INLINE void export_block(const R* I, const PairDev& bp, const PairDev& kp, bool same_pair, double* out) {
    auto tri = [](long long x, long long y) { return x >= y ? x * (x + 1) / 2 + y : y * (y + 1) / 2 + x; };
    const bool same_ab = bp.sa == bp.sb, same_cd = kp.sa == kp.sb;
    sfor<NA>([&](auto p) {
        sfor<NB>([&](auto q) {
            sfor<NC>([&](auto r) {
                sfor<ND>([&](auto u) {
                    const int fi = bp.ao_a + p, fj = bp.ao_b + q, fk = kp.ao_a + r, fl = kp.ao_b + u;
                    const long long ij = tri(fi, fj), kl = tri(fk, fl);
                    if (!((same_ab && fi < fj) || (same_cd && fk < fl) || (same_pair && ij < kl)))
                        out[tri(ij, kl)] = (double)I[((p * NB + q) * NC + r) * ND + u];
                });
            });
        });
    });
}

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
    __shared__ int next_item;
    const bool dscreen = a.dtol > 0.0;
    constexpr bool kDeferJab = NA * NB <= 9;  // s/p bra pairs: J_ab partials in registers
    R jab[kDeferJab ? NA * NB : 1];
    int jab_row = -1, jab_o = 0;  // bra row of the partials, LDS offset of its J_ab block
    // persistent workgroups: each pulls work items from the launch's device queue until it is
    // drained (the counter only grows, so every workgroup leaves the loop)
    for (;;) {
        __syncthreads();  // the previous item's flush has read the tiles and next_item
        if (threadIdx.x == 0) next_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = next_item;
        if (item >= a.nitems) break;
        const WorkItem w = a.items[item];
        const MolDev md = a.mols[w.mol];
        const double* __restrict__ Dm = a.D + md.moff;
        const float* __restrict__ dm = a.dmax + md.dm_off;
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
            rowpre[r + 1] = max(0, min(c, w.c1) - w.c0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            rowpre[0] = 0;
            for (int r = 0; r < nrows; ++r) rowpre[r + 1] += rowpre[r];
        }
        __syncthreads();
        const int total = rowpre[nrows];
        jab_row = -1;
        auto flush_jab = [&]() {
            if constexpr (kDeferJab) {
                if (jab_row >= 0)
                    sfor<NA>([&](auto p) {
                        sfor<NB>([&](auto q) { atomicAdd(&tJab[jab_o + p * fB + q], (ACC)jab[p * NB + q]); });
                    });
            }
        };
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
                    if (density_bound(a.pairs[bpo + i], a.pairs[kpo + j], dm, md.nshell) >= a.dtol)
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
                if (a.eri_out) {  // export instead of digestion (uniform over the launch)
                    export_block<NA, NB, NC, ND, R>(I, bp, kp, same && i == j, a.eri_out);
                    continue;
                }
                if constexpr (kDeferJab) {
                    if (i != jab_row) {
                        flush_jab();
                        jab_row = i;
                        jab_o = bp.slot_a * NA * fB + bp.slot_b * NB;
                        sfor<NA * NB>([&](auto k) { jab[k] = (R)0; });
                    }
                }
                digest<NA, NB, NC, ND, R, ACC, kDeferJab>(I, s, Dm, md.n, bp, kp, tJab, tJcd, tKac, tKad, tKbc, tKbd,
                                                          fB, fC, fD, jab);
            }
            if (dscreen) __syncthreads();  // every thread has read ccount and clist
        }
        flush_jab();
        __syncthreads();
        if (a.eri_out) continue;  // an export has no tiles (the item may have no slab)
        double* out = a.slab + w.slab;
        for (int k = threadIdx.x; k < ntile; k += blockDim.x) out[k] = (double)tiles[k];
    }
}

// Launchers of one class, FP64 and mixed (each specialised in its own translation unit).
template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_jk_class64(const JkArgs& a, int nitems, LaunchShape sh, hipStream_t s);
template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_jk_class32(const JkArgs& a, int nitems, LaunchShape sh, hipStream_t s);

}  // namespace hfg
