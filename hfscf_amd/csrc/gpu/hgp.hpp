// Head-Gordon-Pople evaluation of a contracted shell quartet (ab|cd):
//   1. per primitive quartet: Boys values F_m(T), m = 0..LT, scaled by the quartet prefactor;
//   2. Obara-Saika vertical recursion (VRR) building [e0|f0]^(0) for
//      la <= |e| <= la+lb and lc <= |f| <= lc+ld (the same OS terms hfscf's 1-D recursions
//      apply per Cartesian direction, hfscf: src/integrals/hfscf_osrecur.hpp:211-286);
//   3. contraction over primitives;
//   4. horizontal recursion (HRR) (a, b+1_i| = (a+1_i, b| + AB_i (a, b| on the bra and the same with
//      CD on the ket (hfscf: src/integrals/hfscf_osrecur.hpp:288-312);
//   5. optional Cartesian -> spherical transform (m = -l..l, hfscf: src/math/solid_harmonics.cpp).
// Every loop is unrolled at compile time per angular-momentum class so the intermediates live
// in registers.
#pragma once
#include "cart.hpp"

namespace hfg {

// ---------------------------------------------------------------------------------- VRR
// v[e][f][m] for |e| <= LAB, |f| <= LCD, m <= LT - |e| - |f| (entries outside are unused).
template <int LAB, int LCD, typename R>
struct Vrr {
    static constexpr int LT = LAB + LCD;
    static constexpr int NE = ncum(LAB);
    static constexpr int NF = ncum(LCD);
    R v[NE][NF][LT + 1];

    // F[m] = pfac * F_m(T); PA, WP = W - P, QC, WQ = W - Q; o2p = 1/2p; rp = rho/p;
    // o2q = 1/2q; rq = rho/q; o2pq = 1/2(p+q)
    HFG_INLINE void run(const R* F, const R* PA, const R* WP, const R* QC, const R* WQ, R o2p, R rp, R o2q, R rq,
                        R o2pq) {
        sfor<LT + 1>([&](auto m) { v[0][0][m] = F[m]; });
        sfor<NE - 1>([&](auto e0) {
            constexpr int e = e0 + 1;
            constexpr int le = glvl(e);
            constexpr int d = gdir(e);
            constexpr int e1 = gminus(e, d);
            constexpr int n1 = gexp(e1, d);
            sfor<LT - le + 1>([&](auto m) {
                R t = PA[d] * v[e1][0][m] + WP[d] * v[e1][0][m + 1];
                if constexpr (n1 > 0) {
                    constexpr int e2 = gminus(e1, d);
                    t += (R)n1 * o2p * (v[e2][0][m] - rp * v[e2][0][m + 1]);
                }
                v[e][0][m] = t;
            });
        });
        sfor<NF - 1>([&](auto f0) {
            constexpr int f = f0 + 1;
            constexpr int lf = glvl(f);
            constexpr int d = gdir(f);
            constexpr int f1 = gminus(f, d);
            constexpr int n1 = gexp(f1, d);
            sfor<NE>([&](auto e) {
                constexpr int le = glvl(e);
                constexpr int ne = gexp(e, d);
                sfor<LT - le - lf + 1>([&](auto m) {
                    R t = QC[d] * v[e][f1][m] + WQ[d] * v[e][f1][m + 1];
                    if constexpr (n1 > 0) {
                        constexpr int f2 = gminus(f1, d);
                        t += (R)n1 * o2q * (v[e][f2][m] - rq * v[e][f2][m + 1]);
                    }
                    if constexpr (ne > 0) {
                        constexpr int em = gminus(e, d);
                        t += (R)ne * o2pq * v[em][f1][m + 1];
                    }
                    v[e][f][m] = t;
                });
            });
        });
    }
};

// ---------------------------------------------------------------------------------- HRR
// Contracted [e0|f0] (la <= |e| <= la+lb, lc <= |f| <= lc+ld) -> Cartesian (ab|cd),
// out[((a*nb + b)*nc + c)*nd + d] with a..d indices within their shells.
template <int LA, int LB, int LC, int LD, typename R>
struct Hrr {
    static constexpr int NE0 = ncum(LA - 1), NE = ncum(LA + LB) - NE0;  // e range [NE0, NE0+NE)
    static constexpr int NF0 = ncum(LC - 1), NF = ncum(LC + LD) - NF0;
    static constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);

    // acc[e - NE0][f - NF0] -> out[a][b][c][d]
    HFG_INLINE static void run(const R (&acc)[NE][NF], const R* AB, const R* CD, R* out) {
        // bra: build (a, b| for every f in range
        R bra[na * nb][NF];
        sfor<NF>([&](auto fi) {
            // h[k][e'][b] with |b| = k, |e'| in [LA, LA+LB-k]; stored flat by global indices
            R h[ncum(LA + LB)][ncum(LB)];
            sfor<NE>([&](auto ei) { h[NE0 + ei][0] = acc[ei][fi]; });
            sfor<ncum(LB) - 1>([&](auto b0) {
                constexpr int b = b0 + 1;
                constexpr int lb = glvl(b);
                constexpr int d = gdir(b);
                constexpr int bm = gminus(b, d);
                // a-range for this level: |a| in [LA, LA+LB-lb]
                sfor<ncum(LA + LB - lb) - NE0>([&](auto ai) {
                    constexpr int a = NE0 + ai;
                    constexpr int ap = gplus(a, d);
                    h[a][b] = h[ap][bm] + AB[d] * h[a][bm];
                });
            });
            sfor<na>([&](auto a) {
                sfor<nb>([&](auto b) { bra[a * nb + b][fi] = h[NE0 + a][ncum(LB - 1) + b]; });
            });
        });
        // ket: for every (a, b) build |c, d)
        sfor<na * nb>([&](auto ab) {
            R h[ncum(LC + LD)][ncum(LD)];
            sfor<NF>([&](auto fi) { h[NF0 + fi][0] = bra[ab][fi]; });
            sfor<ncum(LD) - 1>([&](auto d0) {
                constexpr int dd = d0 + 1;
                constexpr int ld = glvl(dd);
                constexpr int dir = gdir(dd);
                constexpr int dm = gminus(dd, dir);
                sfor<ncum(LC + LD - ld) - NF0>([&](auto ci) {
                    constexpr int c = NF0 + ci;
                    constexpr int cp = gplus(c, dir);
                    h[c][dd] = h[cp][dm] + CD[dir] * h[c][dm];
                });
            });
            sfor<nc>([&](auto c) {
                sfor<nd>([&](auto d) { out[(ab * nc + c) * nd + d] = h[NF0 + c][ncum(LD - 1) + d]; });
            });
        });
    }
};

// ---------------------------------------------------------------------------------- spherical
// Cartesian -> spherical coefficient Y_l(c, s) (c Cartesian index, s = l + m), the real solid
// harmonics of x^L-normalised Cartesians (hfscf: src/math/solid_harmonics.cpp:16-109).
// l = 1: (y, z, x); l = 2: (sqrt3 xy, sqrt3 yz, zz - (xx+yy)/2, sqrt3 xz, sqrt3/2 (xx - yy)).
HFG_HD double ylm_coef(int l, int c, int s) {
    if (l == 0) return 1.0;
    if (l == 1) return (s == 0 && c == 1) || (s == 1 && c == 2) || (s == 2 && c == 0) ? 1.0 : 0.0;
    if (l == 2) {
        const double r3 = 1.7320508075688772935;
        switch (s) {
            case 0: return c == 1 ? r3 : 0.0;
            case 1: return c == 4 ? r3 : 0.0;
            case 2: return c == 5 ? 1.0 : (c == 0 || c == 3 ? -0.5 : 0.0);
            case 3: return c == 2 ? r3 : 0.0;
            case 4: return c == 0 ? 0.5 * r3 : (c == 3 ? -0.5 * r3 : 0.0);
        }
    }
    return 0.0;
}

template <bool PURE>
HFG_HD int nfun(int L) {
    return PURE ? 2 * L + 1 : ncart(L);
}

// In-register transform of one index (position POS in 0..3) of a block with dims N0..N3.
template <int L, bool PURE, typename R, int N0, int N1, int N2, int N3, int POS>
struct SphOne {
    static constexpr int NIN[4] = {N0, N1, N2, N3};
    static constexpr int NS = (PURE && L > 0) ? 2 * L + 1 : ncart(L);
    static constexpr int M0 = POS == 0 ? NS : N0, M1 = POS == 1 ? NS : N1, M2 = POS == 2 ? NS : N2,
                         M3 = POS == 3 ? NS : N3;
    HFG_INLINE static void run(const R* in, R* out) {
        sfor<M0>([&](auto i0) {
            sfor<M1>([&](auto i1) {
                sfor<M2>([&](auto i2) {
                    sfor<M3>([&](auto i3) {
                        constexpr int idx[4] = {i0, i1, i2, i3};
                        constexpr int s = idx[POS];
                        R acc = 0;
                        sfor<ncart(L)>([&](auto c) {
                            constexpr double w = ylm_coef(L, c, s);
                            if constexpr (w != 0.0) {
                                constexpr int j0 = POS == 0 ? (int)c : idx[0];
                                constexpr int j1 = POS == 1 ? (int)c : idx[1];
                                constexpr int j2 = POS == 2 ? (int)c : idx[2];
                                constexpr int j3 = POS == 3 ? (int)c : idx[3];
                                if constexpr (w == 1.0)
                                    acc += in[((j0 * N1 + j1) * N2 + j2) * N3 + j3];
                                else
                                    acc += (R)w * in[((j0 * N1 + j1) * N2 + j2) * N3 + j3];
                            }
                        });
                        out[((i0 * M1 + i1) * M2 + i2) * M3 + i3] = acc;
                    });
                });
            });
        });
    }
};

// Full 4-index transform: Cartesian block (ncart) -> output block (nfun); identity when not PURE.
template <int LA, int LB, int LC, int LD, bool PURE, typename R>
struct Sph {
    static constexpr int c0 = ncart(LA), c1 = ncart(LB), c2 = ncart(LC), c3 = ncart(LD);
    static constexpr int s0 = (PURE && LA > 0) ? 2 * LA + 1 : c0, s1 = (PURE && LB > 0) ? 2 * LB + 1 : c1,
                         s2 = (PURE && LC > 0) ? 2 * LC + 1 : c2, s3 = (PURE && LD > 0) ? 2 * LD + 1 : c3;
    static constexpr bool any = PURE && (LA + LB + LC + LD) > 0;
    HFG_INLINE static void run(const R* cart, R* out) {
        if constexpr (!any) {
            sfor<c0 * c1 * c2 * c3>([&](auto i) { out[i] = cart[i]; });
        } else {
            R t1[s0 * c1 * c2 * c3];
            R t2[s0 * s1 * c2 * c3];
            R t3[s0 * s1 * s2 * c3];
            SphOne<LA, PURE, R, c0, c1, c2, c3, 0>::run(cart, t1);
            SphOne<LB, PURE, R, s0, c1, c2, c3, 1>::run(t1, t2);
            SphOne<LC, PURE, R, s0, s1, c2, c3, 2>::run(t2, t3);
            SphOne<LD, PURE, R, s0, s1, s2, c3, 3>::run(t3, out);
        }
    }
};

}  // namespace hfg
