// J/K kernel template on the shell-group lists for the classes with L = la+lb+lc+ld > 0 whose
// pairs hold s shells (one instantiation per class, basis type and precision).
//
// A pair of an s group and a shell (or of two s groups) has up to gc_members member shell pairs
// that share every primitive pair (exponents, centres, 1/2p, P - A) and differ only in K. Each
// primitive quartet of a group quartet gets one prefactor, one set of Boys values (scaled by
// 1/sqrt(p + q) alone) and one vertical recursion, whose results are added into every member
// quartet's accumulators with that member's K_ab K_cd, after hfscf's primitive cutoff on the
// member's own prefactor K_ab K_cd / sqrt(p + q). Each member quartet is then finished as the
// per-shell kernel finishes a quartet: horizontal recursion, element screen, spherical transform
// or Cartesian diagonal, and digestion, after the mean-Schwarz and density screens on its own
// pairs. As in the (ss|ss) group kernel, a work item's `split` lanes share each group quartet over
// slices of the bra primitive pairs and digest their partial integrals.
#pragma once
#include <type_traits>

#include "boys32.hpp"
#include "gc.hpp"
#include "jk_impl.hpp"

namespace hfg {

// The FP32 kernel of the mixed builds at four wavefronts per SIMD (128 VGPRs and 60 bytes of
// scratch per lane for (ps|ss), against 162 VGPRs and three wavefronts unconstrained); the FP64
// kernel unconstrained (forced to three wavefronts it kept 340 bytes per lane in scratch and
// slowed from 8.1 to 11.4 ms on the (H2O)16 build).
template <int LA, int LB, int LC, int LD, bool PURE, typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(std::is_same_v<R, float> ? 4 : 1))) void
jk_gc_kernel(JkArgs a, GcArgs g) {
    constexpr int LT = LA + LB + LC + LD;
    static_assert(LT > 0, "(ss|ss) has its own group kernel");
    constexpr int NA = PURE && LA > 0 ? 2 * LA + 1 : ncart(LA);
    constexpr int NB = PURE && LB > 0 ? 2 * LB + 1 : ncart(LB);
    constexpr int NC = PURE && LC > 0 ? 2 * LC + 1 : ncart(LC);
    constexpr int ND = PURE && LD > 0 ? 2 * LD + 1 : ncart(LD);
    constexpr int MB = gc_members(LA, LB), MK = gc_members(LC, LD);
    using H = Hrr<LA, LB, LC, LD, R>;
    using S = Sph<LA, LB, LC, LD, PURE, R>;
    __shared__ double tiles[kRegions * kBlockFuncs * kBlockFuncs];
    __shared__ int rowpre[kMaxRows + 1];
    __shared__ int next_item;
    const bool dscreen = a.dtol > 0.0;
    R jab[MB][NA * NB];  // J_ab partials of the current bra row, per member
    int jab_o[MB];
    int jab_row = -1;
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
        double* tJab = tiles;
        double* tJcd = tJab + fA * fB;
        double* tKac = tJcd + fC * fD;
        double* tKad = tKac + fA * fC;
        double* tKbc = tKad + fA * fD;
        double* tKbd = tKbc + fB * fC;
        const int ntile = fA * fB + fC * fD + fA * fC + fA * fD + fB * fC + fB * fD;
        for (int k = threadIdx.x; k < ntile; k += blockDim.x) tiles[k] = 0.0;
        const int nrows = w.r1 - w.r0;
        const int bpo = g.goff[w.P], kpo = g.goff[w.Q], kn = g.gcnt[w.Q];
        const bool same = (w.P == w.Q);
        for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
            const int i = w.r0 + r;
            int c = ket_count(g.gkey + kpo, kn, g.gkey[bpo + i], a.tol);
            if (same) c = min(c, i + 1);
            rowpre[r + 1] = max(0, min(c, w.c1) - w.c0);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            rowpre[0] = 0;
            for (int r = 0; r < nrows; ++r) rowpre[r + 1] += rowpre[r];
        }
        __syncthreads();
        // slice-major: neighbouring lanes hold neighbouring group quartets of one slice
        const int split = w.split;
        const int nq = rowpre[nrows];
        const int total = nq * split;
        jab_row = -1;
        auto flush_jab = [&]() {
            if (jab_row >= 0)
                sfor<MB>([&](auto m) {
                    if (jab_o[m] >= 0)
                        sfor<NA>([&](auto p) {
                            sfor<NB>([&](auto q) { atomicAdd(&tJab[jab_o[m] + p * fB + q], (double)jab[m][p * NB + q]); });
                        });
                });
        };
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            const int slice = u / nq, uq = u - slice * nq;
            int i, j;
            locate(uq, rowpre, nrows, w, i, j);
            const GcPairDev gb = g.gpairs[bpo + i];
            const GcPairDev gk = g.gpairs[kpo + j];
            if (i != jab_row) {
                flush_jab();
                jab_row = i;
                sfor<MB>([&](auto m) {
                    sfor<NA * NB>([&](auto k) { jab[m][k] = (R)0; });
                    jab_o[m] = -1;
                    if (m < gb.nmem && gb.mem[m] >= 0) {
                        const PairDev bp = a.pairs[gb.mem[m]];
                        jab_o[m] = bp.slot_a * NA * fB + bp.slot_b * NB;
                    }
                });
            }
            // the group quartet's primitive quartets (this lane's slice of the bra ones)
            R acc[MB][MK][H::NE][H::NF];
            sfor<MB>([&](auto x) {
                sfor<MK>([&](auto y) {
                    sfor<H::NE>([&](auto e) { sfor<H::NF>([&](auto f) { acc[x][y][e][f] = (R)0; }); });
                });
            });
            const int ib1 = (gb.np * (slice + 1)) / split;
            for (int ib = (gb.np * slice) / split; ib < ib1; ++ib) {
                const int ia = gb.off + ib * gb.pstride;
                const double4 a0 = g.gp0[ia];
                const double4 a1 = g.gp1[ia];
                const double4 Kb = g.gK[ia];
                const double Kbd[4] = {Kb.x, Kb.y, Kb.z, Kb.w};
                R Kbm[MB];
                R kbmax = (R)0;
                sfor<MB>([&](auto m) {
                    Kbm[m] = (R)Kbd[m];
                    kbmax = fmax(kbmax, fabs(Kbm[m]));
                });
                for (int jk = 0; jk < gk.np; ++jk) {
                    const int ic = gk.off + jk * gk.pstride;
                    const double4 b0 = g.gp0[ic];
                    const double4 b1 = g.gp1[ic];
                    const double4 Kk = g.gK[ic];
                    const double Kkd[4] = {Kk.x, Kk.y, Kk.z, Kk.w};
                    R Kkm[MK];
                    R kkmax = (R)0;
                    sfor<MK>([&](auto m) {
                        Kkm[m] = (R)Kkd[m];
                        kkmax = fmax(kkmax, fabs(Kkm[m]));
                    });
                    Vrr<LA + LB, LC + LD, R> V;
                    R rs;
                    if constexpr (std::is_same_v<R, float>) {
                        static_assert(LT <= 3, "FP32 Boys values to order 3");
                        // the mixed builds' FP32 primitive quartet (P - Q in FP64 first)
                        const float p = (float)a0.x, q = (float)b0.x;
                        rs = boys32_rsqrt(p + q);
                        // every member's prefactor below the cutoff: the per-shell kernel adds zero
                        if (kbmax * kkmax * rs < (float)a.prim_cut) continue;
                        const float PQx = (float)(a0.y - b0.y), PQy = (float)(a0.z - b0.z), PQz = (float)(a0.w - b0.w);
                        const float r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                        const float ipq = rs * rs;
                        const float T = p * q * ipq * r2;
                        float F[LT + 1];
                        boys32<LT>(T, rs, F);
                        const float rp = q * ipq, rq = p * ipq;
                        const float PA[3] = {(float)a1.x, (float)a1.y, (float)a1.z};
                        const float QC[3] = {(float)b1.x, (float)b1.y, (float)b1.z};
                        const float WP[3] = {-rp * PQx, -rp * PQy, -rp * PQz};
                        const float WQ[3] = {rq * PQx, rq * PQy, rq * PQz};
                        V.run(F, PA, WP, QC, WQ, (float)a1.w, rp, (float)b1.w, rq, 0.5f * ipq);
                    } else {
                        const double p = a0.x, q = b0.x;
                        rs = rsqrt(p + q);
                        if (kbmax * kkmax * rs < a.prim_cut) continue;
                        const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
                        const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                        const double ipq = rs * rs;
                        const double T = p * q * ipq * r2;
                        double F[LT + 1];
                        if constexpr (LT == 1)
                            boys_low<1, double>(T, rs, F);
                        else
                            boys_eval_rs<LT, double>(T, a.boys, rs, F);
                        const double rp = q * ipq, rq = p * ipq;
                        const double PA[3] = {a1.x, a1.y, a1.z};
                        const double QC[3] = {b1.x, b1.y, b1.z};
                        const double WP[3] = {-rp * PQx, -rp * PQy, -rp * PQz};
                        const double WQ[3] = {rq * PQx, rq * PQy, rq * PQz};
                        V.run(F, PA, WP, QC, WQ, a1.w, rp, b1.w, rq, 0.5 * ipq);
                    }
                    sfor<MB>([&](auto x) {
                        if (x < gb.nmem)
                            sfor<MK>([&](auto y) {
                                if (y < gk.nmem) {
                                    const R kk = Kbm[x] * Kkm[y];
                                    const R wgt = (fabs(kk * rs) < (R)a.prim_cut) ? (R)0 : kk;
                                    sfor<H::NE>([&](auto e) {
                                        sfor<H::NF>([&](auto f) {
                                            acc[x][y][e][f] += wgt * V.v[H::NE0 + e][H::NF0 + f][0];
                                        });
                                    });
                                }
                            });
                    });
                }
            }
            // the member quartets, one at a time (rolled loops: one copy of the finishing code;
            // member indices come from memory, the accumulators through select chains, since
            // indexing a register array at run time would place it in scratch)
            const bool self = same && i == j;
            const R AB[3] = {(R)gb.AB[0], (R)gb.AB[1], (R)gb.AB[2]};
            const R CD[3] = {(R)gk.AB[0], (R)gk.AB[1], (R)gk.AB[2]};
#pragma unroll 1
            for (int mb = 0; mb < gb.nmem; ++mb) {
                const int pb = g.gpairs[bpo + i].mem[mb];
                if (pb < 0) continue;
                const PairDev bp = a.pairs[pb];
                const double sb = a.psmean[pb];
#pragma unroll 1
                for (int mk = 0; mk < gk.nmem; ++mk) {
                    const int pk = g.gpairs[kpo + j].mem[mk];
                    if (pk < 0) continue;
                    if (self && mk > mb) continue;  // the group pair with itself: each member quartet once
                    if (sb * a.psmean[pk] < a.tol) continue;  // hfscf's mean-Schwarz shell skip
                    const PairDev kp = a.pairs[pk];
                    if (dscreen && density_bound(bp, kp, dm, md.nshell) < a.dtol) continue;
                    R am[H::NE][H::NF];
                    sfor<H::NE>([&](auto e) {
                        sfor<H::NF>([&](auto f) {
                            R v = (R)0;
                            sfor<MB>([&](auto x) {
                                sfor<MK>([&](auto y) { v = (x == mb && y == mk) ? acc[x][y][e][f] : v; });
                            });
                            am[e][f] = v;
                        });
                    });
                    R I[NA * NB * NC * ND];
                    if constexpr (S::any) {
                        R cart[S::c0 * S::c1 * S::c2 * S::c3];
                        H::run(am, AB, CD, cart);
                        element_screen<LA, LB, LC, LD, R>(bp, kp, a.psq, a.elem_tol, cart);
                        S::run(cart, I);
                    } else {
                        H::run(am, AB, CD, I);
                        element_screen<LA, LB, LC, LD, R>(bp, kp, a.psq, a.elem_tol, I);
                        // hfscf's Cartesian diagonal from the screen matrix, added by the first slice
                        if constexpr (LA == LC && LB == LD) {
                            if (a.elem_tol > 0.0 && bp.sa == kp.sa && bp.sb == kp.sb) {
                                constexpr int nab = ncart(LA) * ncart(LB);
                                const double* s2 = a.psq + bp.sq_off;
                                sfor<nab>([&](auto ab) { I[ab * nab + ab] = slice == 0 ? (R)(s2[ab] * s2[ab]) : (R)0; });
                            }
                        }
                    }
                    double s = 1.0;
                    if (bp.sa == bp.sb) s *= 0.5;
                    if (kp.sa == kp.sb) s *= 0.5;
                    if (self && mk == mb) s *= 0.5;
                    R jl[NA * NB];
                    sfor<NA * NB>([&](auto k) { jl[k] = (R)0; });
                    digest<NA, NB, NC, ND, R, double, true>(I, s, Dm, md.n, bp, kp, tJab, tJcd, tKac, tKad, tKbc, tKbd,
                                                            fB, fC, fD, jl);
                    sfor<MB>([&](auto x) {
                        sfor<NA * NB>([&](auto k) { jab[x][k] += (x == mb) ? jl[k] : (R)0; });
                    });
                }
            }
        }
        flush_jab();
        __syncthreads();
        double* out = a.slab + w.slab;
        for (int k = threadIdx.x; k < ntile; k += blockDim.x) out[k] = tiles[k];
    }
}

// Launchers of one class on the group lists, FP64 and mixed (each in its own translation unit).
template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_jk_gc_class64(const JkArgs& a, const GcArgs& g, int nitems, LaunchShape sh, hipStream_t s);
template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_jk_gc_class32(const JkArgs& a, const GcArgs& g, int nitems, LaunchShape sh, hipStream_t s);

}  // namespace hfg
