// Adjoint two-electron gradient kernel (one instantiation per angular-momentum class).
//
// dE_2e/dX = sum_pqrs G_pqrs (pq|rs)^X with the RHF two-particle density
//   G_pqrs = 2 D_pq D_rs - (D_pr D_qs + D_ps D_qr) / 2
// (the 8-fold symmetrised form of hfscf's sum D_ij D_kl (2 (ij|kl)' - (ik|jl)'),
// hfscf: src/gradient/hfscf_grad_rhf.cpp:121-141). The derivative of a Gaussian with respect
// to its centre is d/dA_x phi_a = 2 alpha phi_{a+1x} - a_x phi_{a-1x}
// (hfscf: src/integrals/hfscf_oserideriv1.cpp:227-352); the D centre follows from
// translational invariance. Per primitive quartet the kernel runs the VRR at L+1, applies
// the HRR to the six shifted classes and contracts them with G at once, so derivative
// integrals are never stored.
#pragma once
#include "boys.hpp"
#include "common.hpp"
#include "hgp.hpp"

namespace hfg {

constexpr int kMaxAtoms = kMaxAtomsPerMolecule;

struct GradArgs {
    const WorkItem* items;
    int nitems;
    const PairDev* pairs;
    const double* psmean;
    const int* bp_off;
    const int* bp_cnt;
    const double4* pp0;
    const double4* pp1;
    const double4* pp2;
    const double* boys;
    const double* D;         // concatenated densities (MolDev::moff, MolDev::n)
    const MolDev* mols;
    const int* shell_atom;   // global shell -> atom within its molecule
    double tol;
    double prim_cut;
    double* slab;            // 3 * natom(mol) partials per work item at WorkItem::gslab
    int* queue;              // work-item counter of this launch (zeroed before the launch)
};

// Spherical -> Cartesian back-transform of one index: out[c] = sum_s Y(c, s) in[s].
template <int L, bool PURE, typename R, int N0, int N1, int N2, int N3, int POS>
struct SphTOne {
    static constexpr int NC = ncart(L);
    static constexpr int M0 = POS == 0 ? NC : N0, M1 = POS == 1 ? NC : N1, M2 = POS == 2 ? NC : N2,
                         M3 = POS == 3 ? NC : N3;
    HFG_INLINE static void run(const R* in, R* out) {
        sfor<M0>([&](auto i0) {
            sfor<M1>([&](auto i1) {
                sfor<M2>([&](auto i2) {
                    sfor<M3>([&](auto i3) {
                        constexpr int idx[4] = {i0, i1, i2, i3};
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
                        out[((i0 * M1 + i1) * M2 + i2) * M3 + i3] = acc;
                    });
                });
            });
        });
    }
};

template <int LA, int LB, int LC, int LD, bool PURE, typename R>
struct SphT {
    static constexpr bool any = PURE && (LA + LB + LC + LD) > 0;
    static constexpr int c0 = ncart(LA), c1 = ncart(LB), c2 = ncart(LC), c3 = ncart(LD);
    static constexpr int s0 = PURE && LA > 0 ? 2 * LA + 1 : c0, s1 = PURE && LB > 0 ? 2 * LB + 1 : c1,
                         s2 = PURE && LC > 0 ? 2 * LC + 1 : c2, s3 = PURE && LD > 0 ? 2 * LD + 1 : c3;
    HFG_INLINE static void run(const R* sph, R* cart) {
        if constexpr (!any) {
            sfor<c0 * c1 * c2 * c3>([&](auto i) { cart[i] = sph[i]; });
        } else {
            R t1[c0 * s1 * s2 * s3];
            R t2[c0 * c1 * s2 * s3];
            R t3[c0 * c1 * c2 * s3];
            if constexpr (LA > 0) SphTOne<LA, PURE, R, s0, s1, s2, s3, 0>::run(sph, t1);
            else sfor<c0 * s1 * s2 * s3>([&](auto i) { t1[i] = sph[i]; });
            if constexpr (LB > 0) SphTOne<LB, PURE, R, c0, s1, s2, s3, 1>::run(t1, t2);
            else sfor<c0 * c1 * s2 * s3>([&](auto i) { t2[i] = t1[i]; });
            if constexpr (LC > 0) SphTOne<LC, PURE, R, c0, c1, s2, s3, 2>::run(t2, t3);
            else sfor<c0 * c1 * c2 * s3>([&](auto i) { t3[i] = t2[i]; });
            if constexpr (LD > 0) SphTOne<LD, PURE, R, c0, c1, c2, s3, 3>::run(t3, cart);
            else sfor<c0 * c1 * c2 * c3>([&](auto i) { cart[i] = t3[i]; });
        }
    }
};

// Contract the HRR output of a class shifted on index POS by +1 or -1 with the Cartesian
// density block G (unshifted class): returns sum over components of
//   (+1) G[..x..] I[..x+1_d..]   or   (-1) x_d G[..x..] I[..x-1_d..]
// for d = 0, 1, 2 in out[d].
template <int LA, int LB, int LC, int LD, int POS, int SHIFT, typename R>
HFG_INLINE void contract_shift(const R* G, const R* I, R* out) {
    constexpr int Ls[4] = {LA, LB, LC, LD};
    constexpr int Lp[4] = {LA + (POS == 0) * SHIFT, LB + (POS == 1) * SHIFT, LC + (POS == 2) * SHIFT,
                           LD + (POS == 3) * SHIFT};
    constexpr int n[4] = {ncart(LA), ncart(LB), ncart(LC), ncart(LD)};
    constexpr int np[4] = {ncart(Lp[0]), ncart(Lp[1]), ncart(Lp[2]), ncart(Lp[3])};
    out[0] = out[1] = out[2] = (R)0;
    sfor<n[0] * n[1] * n[2] * n[3]>([&](auto idx) {
        constexpr int i0 = idx / (n[1] * n[2] * n[3]);
        constexpr int i1 = (idx / (n[2] * n[3])) % n[1];
        constexpr int i2 = (idx / n[3]) % n[2];
        constexpr int i3 = idx % n[3];
        constexpr int ci[4] = {i0, i1, i2, i3};
        constexpr int g = ncum(Ls[POS] - 1) + ci[POS];  // global index of the shifted component
        sfor<3>([&](auto d) {
            constexpr int e = gexp(g, d);
            if constexpr (SHIFT > 0 || e > 0) {
                constexpr int gs = SHIFT > 0 ? gplus(g, d) : gminus(g, d);
                constexpr int cs = lidx(gs);
                constexpr int j0 = POS == 0 ? cs : i0, j1 = POS == 1 ? cs : i1, j2 = POS == 2 ? cs : i2,
                              j3 = POS == 3 ? cs : i3;
                const R v = I[((j0 * np[1] + j1) * np[2] + j2) * np[3] + j3];
                if constexpr (SHIFT > 0)
                    out[d] += G[idx] * v;
                else
                    out[d] += (R)e * G[idx] * v;
            }
        });
    });
}

// Extract the [e0|f0]^(0) range of a shifted class from the VRR table and run its HRR.
template <int LA, int LB, int LC, int LD, int LAB1, int LCD1, typename R>
HFG_INLINE void hrr_from_vrr(const Vrr<LAB1, LCD1, R>& V, const R* AB, const R* CD, R* out) {
    using H = Hrr<LA, LB, LC, LD, R>;
    R acc[H::NE][H::NF];
    sfor<H::NE>([&](auto e) { sfor<H::NF>([&](auto f) { acc[e][f] = V.v[H::NE0 + e][H::NF0 + f][0]; }); });
    H::run(acc, AB, CD, out);
}

// R = double: FP64 throughout. R = float (mixed mode): two-particle density, VRR, HRR and
// contraction in FP32; Boys values and pair data from FP64, per-item atom sums in FP64.
template <int LA, int LB, int LC, int LD, bool PURE, typename R>
__global__ __launch_bounds__(256) void grad_kernel(GradArgs a) {
    constexpr int NA = PURE && LA > 0 ? 2 * LA + 1 : ncart(LA);
    constexpr int NB = PURE && LB > 0 ? 2 * LB + 1 : ncart(LB);
    constexpr int NC = PURE && LC > 0 ? 2 * LC + 1 : ncart(LC);
    constexpr int ND = PURE && LD > 0 ? 2 * LD + 1 : ncart(LD);
    constexpr int ca = ncart(LA), cb = ncart(LB), cc = ncart(LC), cd = ncart(LD);
    constexpr int LT = LA + LB + LC + LD;
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
        int lo = 0, hi = nrows;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rowpre[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        const int i = w.r0 + lo;
        const int j = w.c0 + t - rowpre[lo];
        const PairDev bp = a.pairs[bpo + i];
        const PairDev kp = a.pairs[kpo + j];
        double deg = 8.0;
        if (bp.sa == bp.sb) deg *= 0.5;
        if (kp.sa == kp.sb) deg *= 0.5;
        if (same && i == j) deg *= 0.5;
        // two-particle density in the output basis, then back to Cartesian components
        R G[ca * cb * cc * cd];
        {
            R Gs[NA * NB * NC * ND];
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
            SphT<LA, LB, LC, LD, PURE, R>::run(Gs, G);
        }
        R gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};
        const R AB[3] = {(R)bp.AB[0], (R)bp.AB[1], (R)bp.AB[2]};
        const R CD[3] = {(R)kp.AB[0], (R)kp.AB[1], (R)kp.AB[2]};
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
                const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
                const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                const double ipq = rs * rs;
                const double T = p * q * ipq * r2;
                R F[LT + 2];
                boys_eval<LT + 1, R>(T, a.boys, pfac, F);
                const double rp = q * ipq, rq = p * ipq;
                const R PA[3] = {(R)a1.x, (R)a1.y, (R)a1.z};
                const R QC[3] = {(R)b1.x, (R)b1.y, (R)b1.z};
                const R WP[3] = {(R)(-rp * PQx), (R)(-rp * PQy), (R)(-rp * PQz)};
                const R WQ[3] = {(R)(rq * PQx), (R)(rq * PQy), (R)(rq * PQz)};
                Vrr<LA + LB + 1, LC + LD + 1, R> V;
                V.run(F, PA, WP, QC, WQ, (R)a2.x, (R)rp, (R)b2.x, (R)rq, (R)(0.5 * ipq));
                const R wa = (R)(2.0 * a2.y), wb = (R)(2.0 * a2.z), wc = (R)(2.0 * b2.y);
                R s[3];
                {  // centre A
                    R I[ncart(LA + 1) * cb * cc * cd];
                    hrr_from_vrr<LA + 1, LB, LC, LD>(V, AB, CD, I);
                    contract_shift<LA, LB, LC, LD, 0, 1, R>(G, I, s);
                    sfor<3>([&](auto d) { gA[d] += wa * s[d]; });
                    if constexpr (LA > 0) {
                        R Im[ncart(LA - 1) * cb * cc * cd];
                        hrr_from_vrr<LA - 1, LB, LC, LD>(V, AB, CD, Im);
                        contract_shift<LA, LB, LC, LD, 0, -1, R>(G, Im, s);
                        sfor<3>([&](auto d) { gA[d] -= s[d]; });
                    }
                }
                {  // centre B
                    R I[ca * ncart(LB + 1) * cc * cd];
                    hrr_from_vrr<LA, LB + 1, LC, LD>(V, AB, CD, I);
                    contract_shift<LA, LB, LC, LD, 1, 1, R>(G, I, s);
                    sfor<3>([&](auto d) { gB[d] += wb * s[d]; });
                    if constexpr (LB > 0) {
                        R Im[ca * ncart(LB - 1) * cc * cd];
                        hrr_from_vrr<LA, LB - 1, LC, LD>(V, AB, CD, Im);
                        contract_shift<LA, LB, LC, LD, 1, -1, R>(G, Im, s);
                        sfor<3>([&](auto d) { gB[d] -= s[d]; });
                    }
                }
                {  // centre C
                    R I[ca * cb * ncart(LC + 1) * cd];
                    hrr_from_vrr<LA, LB, LC + 1, LD>(V, AB, CD, I);
                    contract_shift<LA, LB, LC, LD, 2, 1, R>(G, I, s);
                    sfor<3>([&](auto d) { gC[d] += wc * s[d]; });
                    if constexpr (LC > 0) {
                        R Im[ca * cb * ncart(LC - 1) * cd];
                        hrr_from_vrr<LA, LB, LC - 1, LD>(V, AB, CD, Im);
                        contract_shift<LA, LB, LC, LD, 2, -1, R>(G, Im, s);
                        sfor<3>([&](auto d) { gC[d] -= s[d]; });
                    }
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

template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_grad_class64(const GradArgs& a, int nitems, LaunchShape sh, hipStream_t s);
template <int LA, int LB, int LC, int LD, bool PURE>
hipError_t launch_grad_class32(const GradArgs& a, int nitems, LaunchShape sh, hipStream_t s);

hipError_t launch_grad(int la, int lb, int lc, int ld, bool pure, int mixed, const GradArgs& a, int nitems,
                       LaunchShape sh, hipStream_t s);

// Per-molecule gradient sums over work items (fixed order).
struct GradReduceArgs {
    const double* slab;
    const WorkItem* items;     // all work items (gslab offsets)
    const int* mol_items;      // CSR: item indices of each molecule
    const int* mol_items_off;
    const int* out_mol;        // molecule of each output component
    const int* out_k;          // component within the molecule (atom * 3 + direction)
    double* out;               // concatenated natom(mol) x 3 gradients
};
hipError_t launch_grad_reduce(const GradReduceArgs& a, int nout, hipStream_t s);

}  // namespace hfg
