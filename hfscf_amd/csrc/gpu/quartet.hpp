// Contracted shell-quartet evaluation and J/K digestion for one (la lb|lc ld) class.
#pragma once
#include <type_traits>

#include "boys.hpp"
#include "boys32.hpp"
#include "hgp.hpp"
#include "layout.hpp"

namespace hfg {

// Boys values scaled by `scale` as boys_eval, with the large-T branch on one reciprocal square
// root (F_0 = sqrt(pi/4) T^-1/2, 1/T = (T^-1/2)^2) instead of a division and a square root.
template <int M, typename R>
__device__ __forceinline__ void boys_eval_rs(double T, const double* __restrict__ tab, double scale, R* F) {
    static_assert(M <= kBoysMaxM, "Boys order beyond table");
    double f[M + 1];
    if (T < kBoysTmax) {
        const int k = __double2int_rn(T * kBoysStepInv);
        const double dt = (double)k * (1.0 / kBoysStepInv) - T;
        const double* row = tab + k * kBoysCols + M;
        double s = row[6];
        s = row[5] + s * dt * (1.0 / 6.0);
        s = row[4] + s * dt * (1.0 / 5.0);
        s = row[3] + s * dt * (1.0 / 4.0);
        s = row[2] + s * dt * (1.0 / 3.0);
        s = row[1] + s * dt * (1.0 / 2.0);
        s = row[0] + s * dt;
        f[M] = s;
        if constexpr (M > 0) {
            const double eT = exp(-T);
#pragma unroll
            for (int m = M - 1; m >= 0; --m) f[m] = (2.0 * T * f[m + 1] + eT) * (1.0 / (2 * m + 1));
        }
    } else {
        const double r = rsqrt(T);
        f[0] = 0.88622692545275801365 * r;  // sqrt(pi)/2
        if constexpr (M > 0) {
            const double eT = exp(-T);
            const double h = 0.5 * r * r;
#pragma unroll
            for (int m = 0; m < M; ++m) f[m + 1] = ((2 * m + 1) * f[m] - eT) * h;
        }
    }
#pragma unroll
    for (int m = 0; m <= M; ++m) F[m] = (R)(scale * f[m]);
}

// F_0 and F_1 scaled by `scale`, without the table (the (ss|ss) and (ps|ss) classes, whose
// primitive quartets are otherwise dominated by the per-lane table gather):
//   F_0(T) = sqrt(pi)/2 erf(sqrt T) / sqrt T, F_0(0) = 1;
//   F_1(T) = (F_0 - e^-T) / 2T for T >= 0.05, else sum_k (-T)^k / (k! (2k + 3)), k <= 8
// (truncation below 1e-16; relative error against a 50-digit series at most 2.3e-16 for F_0 and
// 5.5e-15 for F_1 over T in [0, 39]).
template <int M, typename R>
__device__ __forceinline__ void boys_low(double T, double scale, R* F) {
    static_assert(M <= 1, "closed form for F_0 and F_1 only");
    double f0 = 1.0, ri = 0.0;
    if (T > 0.0) ri = rsqrt(T);
    // erf(sqrt T) = 1 to within erfc(6) < 2.2e-17 from T = 36 on: the lanes there take F_0 =
    // sqrt(pi/T)/2 directly, and a wavefront whose lanes all have T >= 36 (tight primitives on
    // distant centres) skips the erf evaluation
    if (T >= kBoysTmax)
        f0 = 0.88622692545275801365 * ri;
    else if (T > 0.0)
        f0 = 0.88622692545275801365 * erf(T * ri) * ri;  // sqrt(pi)/2
    F[0] = (R)(scale * f0);
    if constexpr (M == 1) {
        double f1;
        if (T >= 0.05) {
            f1 = (f0 - exp(-T)) * (0.5 * ri * ri);  // 1/T from the square root already taken
        } else {
            f1 = 1.0 / 766080.0;
            f1 = -1.0 / 85680.0 + T * f1;
            f1 = 1.0 / 10800.0 + T * f1;
            f1 = -1.0 / 1560.0 + T * f1;
            f1 = 1.0 / 264.0 + T * f1;
            f1 = -1.0 / 54.0 + T * f1;
            f1 = 1.0 / 14.0 + T * f1;
            f1 = -1.0 / 5.0 + T * f1;
            f1 = 1.0 / 3.0 + T * f1;
        }
        F[1] = (R)(scale * f1);
    }
}

// Contracted Cartesian block (ab|cd) of one shell quartet, out[a][b][c][d] (ncart each).
// prim_cut mirrors hfscf's primitive cutoff |pfac| < integral_tol/1e4
// (hfscf: src/integrals/hfscf_oseri.cpp:73,98-101): a cut primitive quartet contributes zero.
// The next ket primitive's data is loaded before the current one is evaluated, and the cut is
// a multiplication, not a branch, so the loads of one iteration are not split around it.
template <int LA, int LB, int LC, int LD, typename R>
HFG_INLINE void eri_cart(const PairDev& bp, const PairDev& kp, const double4* __restrict__ pp0,
                         const double4* __restrict__ pp1, const double4* __restrict__ pp2,
                         const double* __restrict__ boys_tab, double prim_cut, R* out) {
    constexpr int LT = LA + LB + LC + LD;
    using H = Hrr<LA, LB, LC, LD, R>;
    R acc[H::NE][H::NF];
    sfor<H::NE>([&](auto e) { sfor<H::NF>([&](auto f) { acc[e][f] = (R)0; }); });
    if constexpr (std::is_same_v<R, float> && LT <= 3) {
        // mixed precision, L <= 3: the whole primitive quartet in FP32 (the centre differences
        // P - Q in FP64 first), Boys values from boys32
        const int kn = kp.np, ks = kp.pstride;
        const float cut = (float)prim_cut;
        for (int i = 0; i < bp.np; ++i) {
            const int ia = bp.off + i * bp.pstride;
            const double4 a0 = pp0[ia];
            const double4 a1 = pp1[ia];
            const float o2p = (LT > 0) ? (float)pp2[ia].x : 0.0f;
            const float p = (float)a0.x, Ka = (float)a1.w;
            const float PA[3] = {(float)a1.x, (float)a1.y, (float)a1.z};
            int ib = kp.off;
            double4 b0 = pp0[ib], b1 = pp1[ib];
            double o2q = (LT > 0) ? pp2[ib].x : 0.0;
            for (int j = 0; j < kn; ++j) {
                const int ibn = (j + 1 < kn) ? ib + ks : ib;  // stays in bounds on the last primitive
                const double4 nb0 = pp0[ibn], nb1 = pp1[ibn];
                const double no2q = (LT > 0) ? pp2[ibn].x : 0.0;
                const float q = (float)b0.x;
                const float rs = boys32_rsqrt(p + q);
                float pfac = Ka * (float)b1.w * rs;
                pfac = (fabsf(pfac) < cut) ? 0.0f : pfac;
                const float PQx = (float)(a0.y - b0.y), PQy = (float)(a0.z - b0.z), PQz = (float)(a0.w - b0.w);
                const float r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                const float ipq = rs * rs;
                float T = p * q * ipq * r2;
                float F[LT + 1];
                if constexpr (LT == 0) {
                    // the floor of the FP64 path (F_0 at T >= 1e-8, 1 when r_PQ^2 <= 1e-18)
                    T = (r2 > 1e-18f) ? fmaxf(T, 1e-8f) : 0.0f;
                    boys32<0>(T, pfac, F);
                    acc[0][0] += F[0];
                } else {
                    boys32<LT>(T, pfac, F);
                    const float rp = q * ipq, rq = p * ipq;
                    const float QC[3] = {(float)b1.x, (float)b1.y, (float)b1.z};
                    const float WP[3] = {-rp * PQx, -rp * PQy, -rp * PQz};
                    const float WQ[3] = {rq * PQx, rq * PQy, rq * PQz};
                    Vrr<LA + LB, LC + LD, float> V;
                    V.run(F, PA, WP, QC, WQ, o2p, rp, (float)o2q, rq, 0.5f * ipq);
                    sfor<H::NE>([&](auto e) {
                        sfor<H::NF>([&](auto f) { acc[e][f] += V.v[H::NE0 + e][H::NF0 + f][0]; });
                    });
                }
                b0 = nb0;
                b1 = nb1;
                o2q = no2q;
                ib = ibn;
            }
        }
        const R AB[3] = {(R)bp.AB[0], (R)bp.AB[1], (R)bp.AB[2]};
        const R CD[3] = {(R)kp.AB[0], (R)kp.AB[1], (R)kp.AB[2]};
        H::run(acc, AB, CD, out);
        return;
    }
    const int kn = kp.np, ks = kp.pstride;
    for (int i = 0; i < bp.np; ++i) {
        const int ia = bp.off + i * bp.pstride;
        const double4 a0 = pp0[ia];
        const double4 a1 = pp1[ia];
        const double o2p = pp2[ia].x;
        const double p = a0.x;
        int ib = kp.off;
        double4 b0 = pp0[ib], b1 = pp1[ib];
        double o2q = (LT > 0) ? pp2[ib].x : 0.0;
        for (int j = 0; j < kn; ++j) {
            const int ibn = (j + 1 < kn) ? ib + ks : ib;  // stays in bounds on the last primitive
            const double4 nb0 = pp0[ibn], nb1 = pp1[ibn];
            const double no2q = (LT > 0) ? pp2[ibn].x : 0.0;
            const double q = b0.x;
            const double rs = rsqrt(p + q);
            double pfac = a1.w * b1.w * rs;
            pfac = (fabs(pfac) < prim_cut) ? 0.0 : pfac;
            const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
            const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
            const double ipq = rs * rs;
            const double rho = p * q * ipq;
            double T = rho * r2;
            R F[LT + 1];
            if constexpr (LT == 0) {
                // hfscf floors T at 1e-8 inside F_nu and uses F_n = 1/(2n+1) when r_PQ^2 <= 1e-18
                T = (r2 > 1e-18) ? fmax(T, 1e-8) : 0.0;
                boys_low<0, R>(T, pfac, F);
                acc[0][0] += F[0];
            } else {
                if constexpr (LT == 1)
                    boys_low<1, R>(T, pfac, F);
                else
                    boys_eval_rs<LT, R>(T, boys_tab, pfac, F);
                const double rp = q * ipq, rq = p * ipq;
                const R PA[3] = {(R)a1.x, (R)a1.y, (R)a1.z};
                const R QC[3] = {(R)b1.x, (R)b1.y, (R)b1.z};
                const R WP[3] = {(R)(-rp * PQx), (R)(-rp * PQy), (R)(-rp * PQz)};
                const R WQ[3] = {(R)(rq * PQx), (R)(rq * PQy), (R)(rq * PQz)};
                Vrr<LA + LB, LC + LD, R> V;
                V.run(F, PA, WP, QC, WQ, (R)o2p, (R)rp, (R)o2q, (R)rq, (R)(0.5 * ipq));
                sfor<H::NE>([&](auto e) {
                    sfor<H::NF>([&](auto f) { acc[e][f] += V.v[H::NE0 + e][H::NF0 + f][0]; });
                });
            }
            b0 = nb0;
            b1 = nb1;
            o2q = no2q;
            ib = ibn;
        }
    }
    const R AB[3] = {(R)bp.AB[0], (R)bp.AB[1], (R)bp.AB[2]};
    const R CD[3] = {(R)kp.AB[0], (R)kp.AB[1], (R)kp.AB[2]};
    H::run(acc, AB, CD, out);
}

// hfscf's element screen: a Cartesian component (pq|rs) whose Schwarz bound
// sqrt|(pq|pq)| sqrt|(rs|rs)| is below the threshold is dropped before the spherical transform
// (hfscf: src/integrals/hfscf_oseri.cpp:152-158). sq holds sqrt|(pq|pq)| per pair; the mask is
// only evaluated when the smallest bound of the quartet is below the threshold.
template <int LA, int LB, int LC, int LD, typename R>
HFG_INLINE void element_screen(const PairDev& bp, const PairDev& kp, const double* __restrict__ sq, double tol,
                               R* cart) {
    if (tol <= 0.0 || bp.sq_min * kp.sq_min >= tol) return;
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);
    const double* sb = sq + bp.sq_off;
    const double* sk = sq + kp.sq_off;
    // Written as a multiplication by a 0/1 factor: the equivalent conditional store was
    // mis-generated for the (dd|dp) class on gfx950 (its spill-heavy kernel produced wrong
    // integrals), while this form gives the same results as a CPU build of the same code.
    sfor<na * nb>([&](auto ab) {
        const double x = sb[ab];
        sfor<nc * nd>([&](auto cd) {
            const R keep = (x * sk[cd] >= tol) ? (R)1 : (R)0;
            cart[ab * (nc * nd) + cd] *= keep;
        });
    });
}

// Output-basis block of one quartet (spherical when PURE), out[a][b][c][d].
template <int LA, int LB, int LC, int LD, bool PURE, typename R>
HFG_INLINE void eri_block(const PairDev& bp, const PairDev& kp, const double4* __restrict__ pp0,
                          const double4* __restrict__ pp1, const double4* __restrict__ pp2,
                          const double* __restrict__ boys_tab, double prim_cut, const double* __restrict__ sq,
                          double tol, R* out) {
    using S = Sph<LA, LB, LC, LD, PURE, R>;
    if constexpr (S::any) {
        R cart[S::c0 * S::c1 * S::c2 * S::c3];
        eri_cart<LA, LB, LC, LD, R>(bp, kp, pp0, pp1, pp2, boys_tab, prim_cut, cart);
        element_screen<LA, LB, LC, LD, R>(bp, kp, sq, tol, cart);
        S::run(cart, out);
    } else {
        eri_cart<LA, LB, LC, LD, R>(bp, kp, pp0, pp1, pp2, boys_tab, prim_cut, out);
        element_screen<LA, LB, LC, LD, R>(bp, kp, sq, tol, out);
        // with screening, hfscf stores the diagonal (pq|pq) of a Cartesian basis from its screen
        // matrix (hfscf: src/integrals/hfscf_oseri.cpp:152-158,214-218)
        if constexpr (LA == LC && LB == LD) {
            if (tol > 0.0 && bp.sa == kp.sa && bp.sb == kp.sb) {
                constexpr int nab = ncart(LA) * ncart(LB);
                const double* s2 = sq + bp.sq_off;
                sfor<nab>([&](auto ab) { out[ab * nab + ab] = (R)(s2[ab] * s2[ab]); });
            }
        }
    }
}

// Adds one quartet's contributions into the six LDS tiles. I is the output-basis block,
// s the permutational weight (1/2 for each of a==b, c==d, ab==cd). J tiles receive
// 2 s D I and K tiles s D I; the full matrices are J = T + T^T and K = T + T^T.
// With DEFER_JAB the J_ab contribution is added to the caller's registers jab[NA*NB] instead of
// the LDS tile: the lanes of a wavefront mostly share the bra pair, and their atomics on one
// tile element would serialise.
template <int NA, int NB, int NC, int ND, typename R, typename ACC, bool DEFER_JAB = false>
HFG_INLINE void digest(const R* I, double s, const double* __restrict__ D, int n, const PairDev& bp,
                       const PairDev& kp, ACC* tJab, ACC* tJcd, ACC* tKac, ACC* tKad, ACC* tKbc, ACC* tKbd, int fB,
                       int fC, int fD, R* jab = nullptr) {
    const int ao_a = bp.ao_a, ao_b = bp.ao_b, ao_c = kp.ao_a, ao_d = kp.ao_b;
    const int la0 = bp.slot_a * NA, lb0 = bp.slot_b * NB, lc0 = kp.slot_a * NC, ld0 = kp.slot_b * ND;
    R dab[NA][NB], dcd[NC][ND], dac[NA][NC], dad[NA][ND], dbc[NB][NC], dbd[NB][ND];
    sfor<NA>([&](auto p) {
        sfor<NB>([&](auto q) { dab[p][q] = (R)D[(size_t)(ao_a + p) * n + ao_b + q]; });
        sfor<NC>([&](auto r) { dac[p][r] = (R)D[(size_t)(ao_a + p) * n + ao_c + r]; });
        sfor<ND>([&](auto t) { dad[p][t] = (R)D[(size_t)(ao_a + p) * n + ao_d + t]; });
    });
    sfor<NB>([&](auto q) {
        sfor<NC>([&](auto r) { dbc[q][r] = (R)D[(size_t)(ao_b + q) * n + ao_c + r]; });
        sfor<ND>([&](auto t) { dbd[q][t] = (R)D[(size_t)(ao_b + q) * n + ao_d + t]; });
    });
    sfor<NC>([&](auto r) { sfor<ND>([&](auto t) { dcd[r][t] = (R)D[(size_t)(ao_c + r) * n + ao_d + t]; }); });
    const R sj = (R)(2.0 * s), sk = (R)s;
    // J_ab and K_ac / K_ad
    sfor<NA>([&](auto p) {
        sfor<NB>([&](auto q) {
            R j = 0;
            sfor<NC>([&](auto r) { sfor<ND>([&](auto t) { j += dcd[r][t] * I[((p * NB + q) * NC + r) * ND + t]; }); });
            if constexpr (DEFER_JAB)
                jab[p * NB + q] += sj * j;
            else
                atomicAdd(&tJab[(la0 + p) * fB + lb0 + q], (ACC)(sj * j));
        });
        sfor<NC>([&](auto r) {
            R k = 0;
            sfor<NB>([&](auto q) { sfor<ND>([&](auto t) { k += dbd[q][t] * I[((p * NB + q) * NC + r) * ND + t]; }); });
            atomicAdd(&tKac[(la0 + p) * fC + lc0 + r], (ACC)(sk * k));
        });
        sfor<ND>([&](auto t) {
            R k = 0;
            sfor<NB>([&](auto q) { sfor<NC>([&](auto r) { k += dbc[q][r] * I[((p * NB + q) * NC + r) * ND + t]; }); });
            atomicAdd(&tKad[(la0 + p) * fD + ld0 + t], (ACC)(sk * k));
        });
    });
    // J_cd
    sfor<NC>([&](auto r) {
        sfor<ND>([&](auto t) {
            R j = 0;
            sfor<NA>([&](auto p) { sfor<NB>([&](auto q) { j += dab[p][q] * I[((p * NB + q) * NC + r) * ND + t]; }); });
            atomicAdd(&tJcd[(lc0 + r) * fD + ld0 + t], (ACC)(sj * j));
        });
    });
    // K_bc / K_bd
    sfor<NB>([&](auto q) {
        sfor<NC>([&](auto r) {
            R k = 0;
            sfor<NA>([&](auto p) { sfor<ND>([&](auto t) { k += dad[p][t] * I[((p * NB + q) * NC + r) * ND + t]; }); });
            atomicAdd(&tKbc[(lb0 + q) * fC + lc0 + r], (ACC)(sk * k));
        });
        sfor<ND>([&](auto t) {
            R k = 0;
            sfor<NA>([&](auto p) { sfor<NC>([&](auto r) { k += dac[p][r] * I[((p * NB + q) * NC + r) * ND + t]; }); });
            atomicAdd(&tKbd[(lb0 + q) * fD + ld0 + t], (ACC)(sk * k));
        });
    });
}

}  // namespace hfg
