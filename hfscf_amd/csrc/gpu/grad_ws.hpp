// Workspace form of the adjoint gradient kernel for the classes of highest total angular
// momentum ((dd|dd), (dd|dp), (dp|dd)).
//
// Their recursion tables at L+1 hold thousands of values per lane (7840 doubles for the
// (dd|dd) VRR alone). The register-resident kernel spills them to private (scratch) memory,
// and the runtime refuses a dispatch whose per-lane private segment is that large
// (HSA_STATUS_ERROR_OUT_OF_RESOURCES). Here the tables live in a workspace that the library
// allocates: a fixed grid of kWsLanes lanes, each owning a strided column of the buffer
// (element i of a lane at ws[i * kWsLanes + lane], so the 64 lanes of a wavefront touch 64
// consecutive doubles). The arithmetic is the same as the register kernel's:
//   1. the two-particle density of the quartet, back-transformed to Cartesian components (G);
//   2. per primitive quartet, the VRR [e0|f0]^(m) for |e| <= la+lb+1, |f| <= lc+ld+1,
//      |e|+|f| <= L+1, stored compactly;
//   3. per shifted class (a +- 1_i, b +- 1_i or c +- 1_i), the bra and ket HRR, each final value
//      contracted with G as soon as it is formed, so no shifted integral block is stored.
#pragma once
#include "grad_impl.hpp"

namespace hfg {

constexpr int kWsBlock = 64;
constexpr int kWsBlocks = 512;
constexpr int kWsLanes = kWsBlock * kWsBlocks;  // lanes of one workspace launch

// FP64 classes that take the workspace path (and, in the mixed mode, the (dd|dd) class, whose
// FP32 register kernel is not built).
HFG_HD bool grad_uses_ws(int la, int lb, int lc, int ld) { return la + lb + lc + ld >= 7; }

// Lane-strided view of one region of the workspace.
template <typename R>
struct WsView {
    R* p;  // this lane's element 0 of the region
    HFG_INLINE R& operator[](int i) const { return p[(size_t)i * kWsLanes]; }
    HFG_INLINE WsView at(int off) const { return WsView{p + (size_t)off * kWsLanes}; }
};

// Fence between the steps of the recursions: memory may have changed and every view's pointer
// is a new value. Without it the compiler computes the address of every workspace element
// once (the views are loop-invariant), keeps thousands of them live and spills them to
// private memory, forwards stored values into later reads and hoists reads out of the
// primitive loop: private segments of 12-50 KB per lane, the scratch this kernel exists to
// avoid. It emits no instruction.
template <typename R>
HFG_INLINE void ws_fence_one(WsView<R>& v) {
    asm volatile("" : "+v"(v.p)::"memory");
}
template <typename... V>
HFG_INLINE void ws_fence(V&... v) {
    (ws_fence_one(v), ...);
}

// Compact VRR table: entries (e, f, m) with |e| <= LAB, |f| <= LCD, |e| + |f| <= LTM and
// m <= LTM - |e| - |f|, ordered by e, then f, then m.
template <int LAB, int LCD, int LTM>
struct VrrTab {
    HFG_HD static int nm(int le, int lf) { return LTM - le - lf + 1; }
    HFG_HD static int eblock(int le) {
        int n = 0;
        for (int lf = 0; lf <= LCD && le + lf <= LTM; ++lf) n += ncart(lf) * nm(le, lf);
        return n;
    }
    HFG_HD static int off(int e, int f) {
        const int le = glvl(e), lf = glvl(f);
        int n = 0;
        for (int l = 0; l < le; ++l) n += ncart(l) * eblock(l);
        n += lidx(e) * eblock(le);
        for (int l = 0; l < lf; ++l) n += ncart(l) * nm(le, l);
        return n + lidx(f) * nm(le, lf);
    }
    HFG_HD static int size() {
        int n = 0;
        for (int l = 0; l <= LAB; ++l) n += ncart(l) * eblock(l);
        return n;
    }
};

// VRR into the compact table (the recursion of Vrr::run, limited to |e| + |f| <= LTM). The second
// recursion has a fence after every f (one global-memory round trip per f); FINE puts one after
// every (e, f) block instead, as the (dd|dd) class is built (its build takes two hours and was
// not repeated with the coarser fences).
template <int LAB, int LCD, int LTM, typename R, bool FINE>
HFG_INLINE void vrr_ws(const R* F, const R* PA, const R* WP, const R* QC, const R* WQ, R o2p, R rp, R o2q, R rq,
                       R o2pq, WsView<R> v) {
    using T = VrrTab<LAB, LCD, LTM>;
    constexpr int NE = ncum(LAB), NF = ncum(LCD);
    sfor<LTM + 1>([&](auto m) { v[m] = F[m]; });
    sfor<NE - 1>([&](auto e0) {
        constexpr int e = e0 + 1;
        constexpr int le = glvl(e);
        constexpr int d = gdir(e);
        constexpr int e1 = gminus(e, d);
        constexpr int n1 = gexp(e1, d);
        constexpr int o = T::off(e, 0), o1 = T::off(e1, 0);
        sfor<LTM - le + 1>([&](auto m) {
            R t = PA[d] * v[o1 + m] + WP[d] * v[o1 + m + 1];
            if constexpr (n1 > 0) {
                constexpr int o2 = T::off(gminus(e1, d), 0);
                t += (R)n1 * o2p * (v[o2 + m] - rp * v[o2 + m + 1]);
            }
            v[o + m] = t;
        });
        ws_fence(v);
    });
    sfor<NF - 1>([&](auto f0) {
        constexpr int f = f0 + 1;
        constexpr int lf = glvl(f);
        constexpr int d = gdir(f);
        constexpr int f1 = gminus(f, d);
        constexpr int n1 = gexp(f1, d);
        sfor<NE>([&](auto e) {
            constexpr int le = glvl(e);
            if constexpr (le + lf <= LTM) {
                constexpr int ne = gexp(e, d);
                constexpr int o = T::off(e, f), o1 = T::off(e, f1);
                sfor<LTM - le - lf + 1>([&](auto m) {
                    R t = QC[d] * v[o1 + m] + WQ[d] * v[o1 + m + 1];
                    if constexpr (n1 > 0) {
                        constexpr int o2 = T::off(e, gminus(f1, d));
                        t += (R)n1 * o2q * (v[o2 + m] - rq * v[o2 + m + 1]);
                    }
                    if constexpr (ne > 0) {
                        constexpr int om = T::off(gminus(e, d), f1);
                        t += (R)ne * o2pq * v[om + m + 1];
                    }
                    v[o + m] = t;
                });
                if constexpr (FINE) ws_fence(v);
            }
        });
        if constexpr (!FINE) ws_fence(v);
    });
}

// Sizes of the per-lane regions of one class (in elements).
HFG_HD int ws_shift_bra(int sa, int sb, int sc, int sd) {
    return ncart(sa) * ncart(sb) * (ncum(sc + sd) - ncum(sc - 1));
}
HFG_HD int ws_shift_hb(int sa, int sb) { return (ncum(sa + sb) - ncum(sa - 1)) * ncum(sb); }
HFG_HD int ws_shift_hk(int sc, int sd) { return (ncum(sc + sd) - ncum(sc - 1)) * ncum(sd); }
HFG_HD int ws_max(int a, int b) { return a > b ? a : b; }

struct WsLayout {
    int g, v, bra, hb, hk, total;  // region offsets and the per-lane total
};

HFG_HD int ws_vrr_size(int la, int lb, int lc, int ld) {
    // VrrTab<la+lb+1, lc+ld+1, L+1>::size() with run-time arguments
    const int lab = la + lb + 1, lcd = lc + ld + 1, ltm = la + lb + lc + ld + 1;
    int n = 0;
    for (int le = 0; le <= lab; ++le)
        for (int lf = 0; lf <= lcd && le + lf <= ltm; ++lf) n += ncart(le) * ncart(lf) * (ltm - le - lf + 1);
    return n;
}

HFG_HD WsLayout ws_layout(int la, int lb, int lc, int ld, bool pure) {
    const int ca = ncart(la), cb = ncart(lb), cc = ncart(lc), cd = ncart(ld);
    const int sa = pure && la > 0 ? 2 * la + 1 : ca, sb = pure && lb > 0 ? 2 * lb + 1 : cb,
              sc = pure && lc > 0 ? 2 * lc + 1 : cc, sd = pure && ld > 0 ? 2 * ld + 1 : cd;
    // the spherical density block and the back-transform temporaries share the VRR region
    const int sph = sa * sb * sc * sd + ca * sb * sc * sd + ca * cb * sc * sd + ca * cb * cc * sd;
    int bra = 0, hb = 0, hk = 0;
    for (int pos = 0; pos < 3; ++pos)
        for (int sh = -1; sh <= 1; sh += 2) {
            const int l[4] = {la + (pos == 0) * sh, lb + (pos == 1) * sh, lc + (pos == 2) * sh, ld};
            if (l[pos] < 0) continue;
            bra = ws_max(bra, ws_shift_bra(l[0], l[1], l[2], l[3]));
            hb = ws_max(hb, ws_shift_hb(l[0], l[1]));
            hk = ws_max(hk, ws_shift_hk(l[2], l[3]));
        }
    WsLayout w{};
    w.g = 0;
    w.v = ca * cb * cc * cd;
    w.bra = w.v + ws_max(ws_vrr_size(la, lb, lc, ld), pure ? sph : 0);
    w.hb = w.bra + bra;
    w.hk = w.hb + hb;
    w.total = w.hk + hk;
    return w;
}

// One shifted class (position POS of the unshifted class LA..LD moved by SHIFT = +1 or -1):
// bra HRR for every ket function of the class, then per bra function the ket HRR, each final
// integral I contracted at once with the Cartesian density block G of the unshifted class:
//   SHIFT = +1: out[i] += G[x - 1_i] I[x]   (x_i > 0)
//   SHIFT = -1: out[i] += (x_i + 1) G[x + 1_i] I[x]
// (the sums of contract_shift, taken over the shifted components).
template <int LA, int LB, int LC, int LD, int POS, int SHIFT, typename T, typename R>
HFG_INLINE void shift_contract_ws(WsView<R> vt, WsView<R> G, WsView<R> bra, WsView<R> hb, WsView<R> hk,
                                  const R* AB, const R* CD, R* out) {
    constexpr int SA = LA + (POS == 0) * SHIFT, SB = LB + (POS == 1) * SHIFT, SC = LC + (POS == 2) * SHIFT,
                  SD = LD;
    constexpr int NE0 = ncum(SA - 1), NF0 = ncum(SC - 1), NF = ncum(SC + SD) - NF0;
    constexpr int na = ncart(SA), nb = ncart(SB), nc = ncart(SC);
    constexpr int HB = ncum(SB), HK = ncum(SD);
    constexpr int gn[4] = {ncart(LA), ncart(LB), ncart(LC), ncart(LD)};
    out[0] = out[1] = out[2] = (R)0;
    // bra: (a, b| for every ket function f of the class, bra[(a * nb + b) * NF + f]
    sfor<NF>([&](auto fi) {
        constexpr int f = NF0 + fi;
        if constexpr (SB == 0) {
            sfor<na>([&](auto a) { bra[a * NF + fi] = vt[T::off(NE0 + a, f)]; });
        } else {
            sfor<ncum(SB) - 1>([&](auto b0) {
                constexpr int b = b0 + 1;
                constexpr int lb = glvl(b);
                constexpr int d = gdir(b);
                constexpr int bm = gminus(b, d);
                sfor<ncum(SA + SB - lb) - NE0>([&](auto ai) {
                    constexpr int a = NE0 + ai;
                    constexpr int ap = gplus(a, d);
                    R x1, x0;
                    if constexpr (bm == 0) {
                        x1 = vt[T::off(ap, f)];
                        x0 = vt[T::off(a, f)];
                    } else {
                        x1 = hb[(ap - NE0) * HB + bm];
                        x0 = hb[(a - NE0) * HB + bm];
                    }
                    const R val = x1 + AB[d] * x0;
                    if constexpr (lb == SB)
                        bra[(lidx(a) * nb + lidx(b)) * NF + fi] = val;
                    else
                        hb[(a - NE0) * HB + b] = val;
                });
            });
        }
        ws_fence(vt, G, bra, hb, hk);
    });
    // ket per bra function, contracted with G
    sfor<na * nb>([&](auto ab) {
        ws_fence(vt, G, bra, hb, hk);
        constexpr int ia = ab / nb, ib = ab % nb;
        auto contract = [&](auto kc_, auto kd_, R val) {
            constexpr int kc = decltype(kc_)::value, kd = decltype(kd_)::value;
            constexpr int ci[4] = {ia, ib, kc, kd};
            constexpr int Ls[4] = {SA, SB, SC, SD};
            constexpr int g = ncum(Ls[POS] - 1) + ci[POS];
            sfor<3>([&](auto x) {
                constexpr int ex = gexp(g, x);
                if constexpr (SHIFT < 0 || ex > 0) {
                    constexpr int gu = SHIFT > 0 ? gminus(g, x) : gplus(g, x);
                    constexpr int cu = lidx(gu);
                    constexpr int j0 = POS == 0 ? cu : ia, j1 = POS == 1 ? cu : ib, j2 = POS == 2 ? cu : kc;
                    constexpr int gi = ((j0 * gn[1] + j1) * gn[2] + j2) * gn[3] + kd;
                    if constexpr (SHIFT > 0)
                        out[x] += G[gi] * val;
                    else
                        out[x] += (R)(ex + 1) * G[gi] * val;
                }
            });
        };
        if constexpr (SD == 0) {
            sfor<nc>([&](auto c) { contract(c, ic<0>{}, bra[ab * NF + c]); });
        } else {
            sfor<ncum(SD) - 1>([&](auto d0) {
                constexpr int dd = d0 + 1;
                constexpr int ld = glvl(dd);
                constexpr int dir = gdir(dd);
                constexpr int dm = gminus(dd, dir);
                sfor<ncum(SC + SD - ld) - NF0>([&](auto ci) {
                    constexpr int c = NF0 + ci;
                    constexpr int cp = gplus(c, dir);
                    R x1, x0;
                    if constexpr (dm == 0) {
                        x1 = bra[ab * NF + (cp - NF0)];
                        x0 = bra[ab * NF + (c - NF0)];
                    } else {
                        x1 = hk[(cp - NF0) * HK + dm];
                        x0 = hk[(c - NF0) * HK + dm];
                    }
                    const R val = x1 + CD[dir] * x0;
                    if constexpr (ld == SD)
                        contract(ic<lidx(c)>{}, ic<lidx(dd)>{}, val);
                    else
                        hk[(c - NF0) * HK + dd] = val;
                });
            });
        }
    });
}

// Spherical -> Cartesian back-transform of index POS of a block with dims N0..N3 (the
// transform of SphTOne, between workspace regions): out[c] = sum_s Y(c, s) in[s].
template <int L, int N0, int N1, int N2, int N3, int POS, typename R>
HFG_INLINE void sph_back_one_ws(WsView<R> in, WsView<R> out) {
    constexpr int NC = ncart(L);
    constexpr int M0 = POS == 0 ? NC : N0, M1 = POS == 1 ? NC : N1, M2 = POS == 2 ? NC : N2,
                  M3 = POS == 3 ? NC : N3;
    sfor<M0 * M1 * M2 * M3>([&](auto o) {
        constexpr int idx[4] = {o / (M1 * M2 * M3), (o / (M2 * M3)) % M1, (o / M3) % M2, o % M3};
        constexpr int c = idx[POS];
        R acc = 0;
        sfor<2 * L + 1>([&](auto s) {
            constexpr double w = ylm_coef(L, c, s);
            if constexpr (w != 0.0) {
                constexpr int j0 = POS == 0 ? (int)s : idx[0];
                constexpr int j1 = POS == 1 ? (int)s : idx[1];
                constexpr int j2 = POS == 2 ? (int)s : idx[2];
                constexpr int j3 = POS == 3 ? (int)s : idx[3];
                acc += (R)w * in[((j0 * N1 + j1) * N2 + j2) * N3 + j3];
            }
        });
        out[o] = acc;
    });
}

// Spherical -> Cartesian back-transform of the density block with its temporaries in the
// workspace (the steps of SphT).
template <int LA, int LB, int LC, int LD, typename R>
HFG_INLINE void sph_back_ws(WsView<R> sph, WsView<R> t1, WsView<R> t2, WsView<R> t3, WsView<R> cart) {
    constexpr int c0 = ncart(LA), c1 = ncart(LB), c2 = ncart(LC), c3 = ncart(LD);
    constexpr int s0 = LA > 0 ? 2 * LA + 1 : 1, s1 = LB > 0 ? 2 * LB + 1 : 1, s2 = LC > 0 ? 2 * LC + 1 : 1,
                  s3 = LD > 0 ? 2 * LD + 1 : 1;
    if constexpr (LA > 0) sph_back_one_ws<LA, s0, s1, s2, s3, 0, R>(sph, t1);
    else sfor<c0 * s1 * s2 * s3>([&](auto i) { t1[i] = sph[i]; });
    if constexpr (LB > 0) sph_back_one_ws<LB, c0, s1, s2, s3, 1, R>(t1, t2);
    else sfor<c0 * c1 * s2 * s3>([&](auto i) { t2[i] = t1[i]; });
    if constexpr (LC > 0) sph_back_one_ws<LC, c0, c1, s2, s3, 2, R>(t2, t3);
    else sfor<c0 * c1 * c2 * s3>([&](auto i) { t3[i] = t2[i]; });
    if constexpr (LD > 0) sph_back_one_ws<LD, c0, c1, c2, s3, 3, R>(t3, cart);
    else sfor<c0 * c1 * c2 * c3>([&](auto i) { cart[i] = t3[i]; });
}

// FP64 workspace gradient kernel: the work distribution, screening, weights and per-item
// atom sums of grad_kernel; launched as exactly kWsBlocks x kWsBlock lanes.
template <int LA, int LB, int LC, int LD, bool PURE>
__global__ __launch_bounds__(kWsBlock) void grad_ws_kernel(GradArgs a, double* __restrict__ ws) {
    using R = double;
    constexpr int NA = PURE && LA > 0 ? 2 * LA + 1 : ncart(LA);
    constexpr int NB = PURE && LB > 0 ? 2 * LB + 1 : ncart(LB);
    constexpr int NC = PURE && LC > 0 ? 2 * LC + 1 : ncart(LC);
    constexpr int ND = PURE && LD > 0 ? 2 * LD + 1 : ncart(LD);
    constexpr int ca = ncart(LA), cb = ncart(LB);
    constexpr int LT = LA + LB + LC + LD;
    using T = VrrTab<LA + LB + 1, LC + LD + 1, LT + 1>;
    constexpr WsLayout L = ws_layout(LA, LB, LC, LD, PURE);
    const WsView<R> W{ws + blockIdx.x * kWsBlock + threadIdx.x};
    WsView<R> G = W.at(L.g), V = W.at(L.v), Bra = W.at(L.bra), Hb = W.at(L.hb), Hk = W.at(L.hk);
    __shared__ double gsum[kMaxAtoms * 3];
    __shared__ int rowpre[kMaxRows + 1];
    __shared__ int next_item;
    for (;;) {
        __syncthreads();  // the previous item's flush has read gsum and next_item
        if (threadIdx.x == 0) next_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = next_item;
        if (item >= a.nitems) break;
        const WorkItem w = a.items[item];
        const MolDev md = a.mols[w.mol];
        const double* __restrict__ Dm = a.D + md.moff;
        const int natom3 = md.natom * 3;
        for (int k = threadIdx.x; k < natom3; k += blockDim.x) gsum[k] = 0.0;
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
        for (int t = threadIdx.x; t < total; t += blockDim.x) {
            int i, j;
            locate(t, rowpre, nrows, w, i, j);
            const PairDev bp = a.pairs[bpo + i];
            const PairDev kp = a.pairs[kpo + j];
            double deg = 8.0;
            if (bp.sa == bp.sb) deg *= 0.5;
            if (kp.sa == kp.sb) deg *= 0.5;
            if (same && i == j) deg *= 0.5;
            ws_fence(G, V);
            {  // two-particle density of the quartet, then Cartesian components in G
                const WsView<R> Gs = PURE ? V : G;
                const int oa = bp.ao_a, ob = bp.ao_b, oc = kp.ao_a, od = kp.ao_b;
                const int n = md.n;
                sfor<NA>([&](auto p) {
                    sfor<NB>([&](auto q) {
                        const R dpq = Dm[(size_t)(oa + p) * n + ob + q];
                        sfor<NC>([&](auto r) {
                            const R dpr = Dm[(size_t)(oa + p) * n + oc + r];
                            const R dqr = Dm[(size_t)(ob + q) * n + oc + r];
                            sfor<ND>([&](auto s) {
                                const R drs = Dm[(size_t)(oc + r) * n + od + s];
                                const R dqs = Dm[(size_t)(ob + q) * n + od + s];
                                const R dps = Dm[(size_t)(oa + p) * n + od + s];
                                Gs[((p * NB + q) * NC + r) * ND + s] =
                                    (R)deg * ((R)2 * dpq * drs - (R)0.5 * (dpr * dqs + dps * dqr));
                            });
                        });
                    });
                });
                if constexpr (PURE) {
                    constexpr int o1 = NA * NB * NC * ND, o2 = o1 + ca * NB * NC * ND, o3 = o2 + ca * cb * NC * ND;
                    sph_back_ws<LA, LB, LC, LD, R>(V, V.at(o1), V.at(o2), V.at(o3), G);
                }
            }
            R gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
            const R AB[3] = {bp.AB[0], bp.AB[1], bp.AB[2]};
            const R CD[3] = {kp.AB[0], kp.AB[1], kp.AB[2]};
            for (int ip = 0; ip < bp.np; ++ip) {
                const int ia = bp.off + ip * bp.pstride;
                const double4 a0 = a.pp0[ia];
                const double4 a1 = a.pp1[ia];
                const double4 a2 = a.pp2[ia];
                const double p = a0.x;
                for (int jp = 0; jp < kp.np; ++jp) {
                    const int ib = kp.off + jp * kp.pstride;
                    const double4 b0 = a.pp0[ib];
                    const double4 b1 = a.pp1[ib];
                    const double4 b2 = a.pp2[ib];
                    const double q = b0.x;
                    const double rs = rsqrt(p + q);
                    const double pfac = a1.w * b1.w * rs;
                    if (fabs(pfac) < a.prim_cut) continue;
                    ws_fence(G, V, Bra, Hb, Hk);
                    const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
                    const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                    const double ipq = rs * rs;
                    const double Tb = p * q * ipq * r2;
                    R F[LT + 2];
                    boys_eval<LT + 1, R>(Tb, a.boys, pfac, F);
                    const double rp = q * ipq, rq = p * ipq;
                    const R PA[3] = {a1.x, a1.y, a1.z};
                    const R QC[3] = {b1.x, b1.y, b1.z};
                    const R WP[3] = {-rp * PQx, -rp * PQy, -rp * PQz};
                    const R WQ[3] = {rq * PQx, rq * PQy, rq * PQz};
                    vrr_ws<LA + LB + 1, LC + LD + 1, LT + 1, R, (LT >= 8)>(F, PA, WP, QC, WQ, a2.x, rp, b2.x, rq, 0.5 * ipq, V);
                    const R wa = 2.0 * a2.y, wb = 2.0 * a2.z, wc = 2.0 * b2.y;
                    R s[3];
                    shift_contract_ws<LA, LB, LC, LD, 0, 1, T, R>(V, G, Bra, Hb, Hk, AB, CD, s);
                    sfor<3>([&](auto d) { gA[d] += wa * s[d]; });
                    if constexpr (LA > 0) {
                        shift_contract_ws<LA, LB, LC, LD, 0, -1, T, R>(V, G, Bra, Hb, Hk, AB, CD, s);
                        sfor<3>([&](auto d) { gA[d] -= s[d]; });
                    }
                    shift_contract_ws<LA, LB, LC, LD, 1, 1, T, R>(V, G, Bra, Hb, Hk, AB, CD, s);
                    sfor<3>([&](auto d) { gB[d] += wb * s[d]; });
                    if constexpr (LB > 0) {
                        shift_contract_ws<LA, LB, LC, LD, 1, -1, T, R>(V, G, Bra, Hb, Hk, AB, CD, s);
                        sfor<3>([&](auto d) { gB[d] -= s[d]; });
                    }
                    shift_contract_ws<LA, LB, LC, LD, 2, 1, T, R>(V, G, Bra, Hb, Hk, AB, CD, s);
                    sfor<3>([&](auto d) { gC[d] += wc * s[d]; });
                    if constexpr (LC > 0) {
                        shift_contract_ws<LA, LB, LC, LD, 2, -1, T, R>(V, G, Bra, Hb, Hk, AB, CD, s);
                        sfor<3>([&](auto d) { gC[d] -= s[d]; });
                    }
                }
            }
            const int* sat = a.shell_atom + md.shell0;
            const int atA = sat[bp.sa], atB = sat[bp.sb];
            const int atC = sat[kp.sa], atD = sat[kp.sb];
            for (int d = 0; d < 3; ++d) {
                const double xa = gA[d], xb = gB[d], xc = gC[d];
                atomicAdd(&gsum[atA * 3 + d], xa);
                atomicAdd(&gsum[atB * 3 + d], xb);
                atomicAdd(&gsum[atC * 3 + d], xc);
                atomicAdd(&gsum[atD * 3 + d], -(xa + xb + xc));
            }
        }
        __syncthreads();
        double* out = a.slab + w.gslab;
        for (int k = threadIdx.x; k < natom3; k += blockDim.x) out[k] = gsum[k];
    }
}

// Launchers of the workspace classes (one object per class and basis type).
template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_grad_ws(const GradArgs& a, double* ws, int nitems, hipStream_t s);

// Doubles of workspace a class launch needs (0 for the classes on the register kernel).
size_t grad_ws_doubles(int la, int lb, int lc, int ld, bool pure, int mixed);

// Gradient launch of one class; ws: workspace of at least grad_ws_doubles(...) doubles.
hipError_t launch_grad(int la, int lb, int lc, int ld, bool pure, int mixed, const GradArgs& a, double* ws,
                       int nitems, LaunchShape sh, hipStream_t s);

}  // namespace hfg
