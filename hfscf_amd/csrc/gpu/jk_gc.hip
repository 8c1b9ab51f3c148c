// (ss|ss) J/K (FP64, and FP32 for the mixed builds) on the lists of shell groups: the shells of one atom and angular momentum with
// identical exponents (cc-pVDZ oxygen has two s shells on the same nine exponents) share their
// primitive pairs, so each primitive quartet of a group quartet is evaluated once (prefactor,
// Boys value) and contracted into every member shell quartet with that member's coefficients.
// Every member quartet then gets what the per-shell kernel gives it: hfscf's primitive cutoff on
// its own prefactor, the mean-Schwarz shell skip and the density screen on its own pairs, the
// element screen, the Cartesian diagonal from the screen matrix, and its digestion.
// A group quartet of many primitive quartets (two 9-primitive groups: 6561) would hold one lane
// for as long as its members take on many lanes of the per-shell kernel, so the work item's
// `split` lanes share it, each over a slice of the bra primitive pairs; each lane digests its
// partial integrals (the digestion and the element screen are linear in the integral value, and
// the diagonal value that replaces a computed one is added by the first slice only).
#include <type_traits>

#include "boys32.hpp"
#include "gc.hpp"
#include "jk_gc_impl.hpp"
#include "jk_impl.hpp"

namespace hfg {

// Three wavefronts per SIMD: the FP64 kernel then holds 168 VGPRs and keeps 84 bytes per lane in
// scratch, against 185 VGPRs and two wavefronts without the attribute; alone on the (H2O)16 /
// cc-pVDZ build it took 5.0 ms instead of 5.9 (the FP32 kernel fits four wavefronts either way).
// The FP64 (ps|ss) group kernel, forced the same way, spilled 340 bytes per lane and slowed from
// 8.1 to 11.4 ms.
template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void jk_gc_ssss_kernel(JkArgs a, GcArgs g) {
    __shared__ double tiles[kRegions * kBlockFuncs * kBlockFuncs];
    __shared__ int rowpre[kMaxRows + 1];
    __shared__ int next_item;
    const bool dscreen = a.dtol > 0.0;
    const R cut = (R)a.prim_cut;
    R jab[kGcMaxShells * kGcMaxShells];  // J_ab partials of the current bra row, per member
    int jab_o[kGcMaxShells * kGcMaxShells];
    int jab_row = -1, jab_n = 0;
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
        // each group quartet on `split` lanes, each over a slice of the bra primitive pairs
        const int split = w.split;
        const int total = rowpre[nrows] * split;
        jab_row = -1;
        auto flush_jab = [&]() {
            if (jab_row >= 0)
#pragma unroll
                for (int m = 0; m < kGcMaxMembers; ++m)
                    if (m < jab_n && jab_o[m] >= 0) atomicAdd(&tJab[jab_o[m]], jab[m]);
        };
        for (int u = threadIdx.x; u < total; u += blockDim.x) {
            const int uq = u / split, slice = u - uq * split;
            int i, j;
            locate(uq, rowpre, nrows, w, i, j);
            const GcPairDev gb = g.gpairs[bpo + i];
            const GcPairDev gk = g.gpairs[kpo + j];
            if (i != jab_row) {
                flush_jab();
                jab_row = i;
                jab_n = gb.nmem;
                for (int m = 0; m < kGcMaxMembers; ++m) {
                    jab[m] = (R)0;
                    jab_o[m] = -1;
                    if (m < gb.nmem && gb.mem[m] >= 0) {
                        const PairDev bp = a.pairs[gb.mem[m]];
                        jab_o[m] = bp.slot_a * fB + bp.slot_b;
                    }
                }
            }
            // the group quartet's primitive quartets, contracted into every member quartet
            R acc[kGcMaxMembers][kGcMaxMembers];
#pragma unroll
            for (int mb = 0; mb < kGcMaxMembers; ++mb)
#pragma unroll
                for (int mk = 0; mk < kGcMaxMembers; ++mk) acc[mb][mk] = (R)0;
            const int ib1 = (gb.np * (slice + 1)) / split;
            for (int ib = (gb.np * slice) / split; ib < ib1; ++ib) {
                const double4 a0 = g.gp0[gb.off + ib * gb.pstride];
                const double4 Kb = g.gK[gb.off + ib * gb.pstride];
                const R Kbm[4] = {(R)Kb.x, (R)Kb.y, (R)Kb.z, (R)Kb.w};
                const R kbmax = fmax(fmax(fabs(Kbm[0]), fabs(Kbm[1])), fmax(fabs(Kbm[2]), fabs(Kbm[3])));
                for (int jk = 0; jk < gk.np; ++jk) {
                    const double4 b0 = g.gp0[gk.off + jk * gk.pstride];
                    const double4 Kk = g.gK[gk.off + jk * gk.pstride];
                    const R Kkm[4] = {(R)Kk.x, (R)Kk.y, (R)Kk.z, (R)Kk.w};
                    const R kkmax = fmax(fmax(fabs(Kkm[0]), fabs(Kkm[1])), fmax(fabs(Kkm[2]), fabs(Kkm[3])));
                    R rs, F0[1];
                    if constexpr (std::is_same_v<R, float>) {
                        rs = boys32_rsqrt((float)a0.x + (float)b0.x);
                    } else {
                        rs = rsqrt(a0.x + b0.x);
                    }
                    // every member's prefactor |K_ab K_cd| / sqrt(p + q) is below the cutoff (the
                    // products are monotone in their factors): the per-shell kernel adds zero
                    if (kbmax * kkmax * rs < cut) continue;
                    if constexpr (std::is_same_v<R, float>) {
                        // the mixed builds' FP32 primitive quartet (P - Q in FP64 first)
                        const float PQx = (float)(a0.y - b0.y), PQy = (float)(a0.z - b0.z), PQz = (float)(a0.w - b0.w);
                        const float r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                        float T = (float)a0.x * (float)b0.x * (rs * rs) * r2;
                        T = (r2 > 1e-18f) ? fmaxf(T, 1e-8f) : 0.0f;
                        boys32<0>(T, 1.0f, F0);
                    } else {
                        const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
                        const double r2 = PQx * PQx + PQy * PQy + PQz * PQz;
                        const double ipq = rs * rs;
                        double T = a0.x * b0.x * ipq * r2;
                        // hfscf floors T at 1e-8 inside F_nu and uses F_0 = 1 when r_PQ^2 <= 1e-18
                        T = (r2 > 1e-18) ? fmax(T, 1e-8) : 0.0;
                        boys_low<0, double>(T, 1.0, F0);
                    }
#pragma unroll
                    for (int mb = 0; mb < kGcMaxMembers; ++mb) {
                        if (mb >= gb.nmem) continue;
                        const R kb = Kbm[mb];
#pragma unroll
                        for (int mk = 0; mk < kGcMaxMembers; ++mk) {
                            if (mk >= gk.nmem) continue;
                            R pfac = kb * Kkm[mk] * rs;
                            pfac = (fabs(pfac) < cut) ? (R)0 : pfac;
                            acc[mb][mk] += pfac * F0[0];
                        }
                    }
                }
            }
            // the member quartets, one at a time (the loops stay rolled: one copy of the digestion;
            // member indices come from memory and the register arrays are read through select
            // chains, since indexing a register array at run time would place it in scratch)
            const bool self = same && i == j;
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
                    if (self && mk > mb) continue;  // the group paired with itself: each member quartet once
                    if (sb * a.psmean[pk] < a.tol) continue;  // hfscf's mean-Schwarz shell skip
                    const PairDev kp = a.pairs[pk];
                    if (dscreen && density_bound(bp, kp, dm, md.nshell) < a.dtol) continue;
                    R I[1] = {(R)0};
#pragma unroll
                    for (int x = 0; x < kGcMaxMembers; ++x)
#pragma unroll
                        for (int y = 0; y < kGcMaxMembers; ++y) I[0] = (x == mb && y == mk) ? acc[x][y] : I[0];
                    element_screen<0, 0, 0, 0, R>(bp, kp, a.psq, a.elem_tol, I);
                    if (a.elem_tol > 0.0 && bp.sa == kp.sa && bp.sb == kp.sb) {
                        // hfscf's Cartesian diagonal from the screen matrix, added by the first slice
                        const double s2 = a.psq[bp.sq_off];
                        I[0] = slice == 0 ? (R)(s2 * s2) : (R)0;
                    }
                    double s = 1.0;
                    if (bp.sa == bp.sb) s *= 0.5;
                    if (kp.sa == kp.sb) s *= 0.5;
                    if (self && mk == mb) s *= 0.5;
                    R jl[1] = {(R)0};
                    digest<1, 1, 1, 1, R, double, true>(I, s, Dm, md.n, bp, kp, tJab, tJcd, tKac, tKad, tKbc,
                                                             tKbd, fB, fC, fD, jl);
#pragma unroll
                    for (int x = 0; x < kGcMaxMembers; ++x) jab[x] += (x == mb) ? jl[0] : (R)0;
                }
            }
        }
        flush_jab();
        __syncthreads();
        double* out = a.slab + w.slab;
        for (int k = threadIdx.x; k < ntile; k += blockDim.x) out[k] = tiles[k];
    }
}

#define HFG_GC_CLASS(A, B, C, D, P)                                                                    \
    template <>                                                                                        \
    hipError_t launch_jk_gc_class64<A, B, C, D, P>(const JkArgs&, const GcArgs&, int, LaunchShape, hipStream_t); \
    template <>                                                                                        \
    hipError_t launch_jk_gc_class32<A, B, C, D, P>(const JkArgs&, const GcArgs&, int, LaunchShape, hipStream_t);
#define HFG_GC_CLASSES(X) X(1, 0, 0, 0)
#define HFG_GC_DECL(A, B, C, D) HFG_GC_CLASS(A, B, C, D, true) HFG_GC_CLASS(A, B, C, D, false)
HFG_GC_CLASSES(HFG_GC_DECL)

bool jk_gc_available(int la, int lb, int lc, int ld, bool) {
    const int key = la * 1000 + lb * 100 + lc * 10 + ld;
#define HFG_GC_KEY(A, B, C, D) if (key == A * 1000 + B * 100 + C * 10 + D) return true;
    HFG_GC_CLASSES(HFG_GC_KEY)
#undef HFG_GC_KEY
    return key == 0;
}

hipError_t launch_jk_gc(int la, int lb, int lc, int ld, bool pure, int mixed, const JkArgs& a, const GcArgs& g,
                        int nitems, LaunchShape sh, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    const int key = la * 1000 + lb * 100 + lc * 10 + ld;
    if (key == 0) {
        JkArgs b = a;
        b.nitems = nitems;
        const int block = sh.block > 0 ? sh.block : 256;
        const int grid = sh.grid > 0 ? std::min(sh.grid, nitems) : nitems;
        if (mixed)
            hipLaunchKernelGGL(jk_gc_ssss_kernel<float>, dim3(grid), dim3(block), 0, s, b, g);
        else
            hipLaunchKernelGGL(jk_gc_ssss_kernel<double>, dim3(grid), dim3(block), 0, s, b, g);
        return hipGetLastError();
    }
#define HFG_GC_CALL(A, B, C, D)                                                                           \
    if (key == A * 1000 + B * 100 + C * 10 + D) {                                                         \
        if (pure)                                                                                         \
            return mixed ? launch_jk_gc_class32<A, B, C, D, true>(a, g, nitems, sh, s)                   \
                         : launch_jk_gc_class64<A, B, C, D, true>(a, g, nitems, sh, s);                  \
        return mixed ? launch_jk_gc_class32<A, B, C, D, false>(a, g, nitems, sh, s)                      \
                     : launch_jk_gc_class64<A, B, C, D, false>(a, g, nitems, sh, s);                     \
    }
    HFG_GC_CLASSES(HFG_GC_CALL)
#undef HFG_GC_CALL
    return hipErrorInvalidValue;
}

}  // namespace hfg
