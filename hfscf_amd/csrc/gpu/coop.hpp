// Wavefront-cooperative J/K kernel for the classes of highest angular momentum.
//
// The class kernels of jk_impl.hpp give each lane a whole shell quartet: every recursion value is
// a register of that lane. For (dd|dd), (dd|dp) and (dp|dd) the tables hold thousands of values
// (3,675 VRR entries for (dd|dd)), the kernels spill 9-14 KB per lane to scratch, and a launch is
// a few hundred lanes each running a long serial chain: latency-bound, with most of the chip idle.
// Here one workgroup evaluates one quartet at a time and every step is spread over its lanes,
// with the tables in LDS:
//   1. Boys values F_m(T), one order per lane (Taylor step of the table at that order, or the
//      upward recursion beyond the table), scaled by the quartet prefactor;
//   2. Obara-Saika VRR [e0|f0]^(m) level by level (|e| for the bra, then |f| for the ket): the
//      targets of one level depend only on lower levels, so the lanes share a level's targets;
//   3. contraction over primitive quartets, each lane holding its share of [e0|f0]^(0) in
//      registers;
//   4. bra HRR per ket function and ket HRR per bra function pair, level by level over |b| and |d|;
//   5. hfscf's element screen and Cartesian diagonal, the spherical transform one index at a time;
//   6. digestion: each lane owns some elements of the six J/K tiles and sums its products.
// The recursion steps are run from programs (lists of targets and sources per level) that the
// host generates per class (coop.hip), so one kernel per precision serves every class and the
// build compiles in seconds. Work items, screening, tiles and slab flushes are those of jk_kernel.
#pragma once
#include <cmath>
#include <vector>

#include "coop_args.hpp"
#include "jk_impl.hpp"

namespace hfg {

// F_k(T) (one order per lane): the Taylor step of boys_eval_rs at order k for T < kBoysTmax,
// the upward recursion from F_0 = sqrt(pi/4T) otherwise.
__host__ __device__ inline double boys_order(int k, double T, const double* __restrict__ tab) {
    if (T < kBoysTmax) {
#if defined(__HIP_DEVICE_COMPILE__)
        const int g = __double2int_rn(T * kBoysStepInv);
#else
        const int g = (int)std::nearbyint(T * kBoysStepInv);
#endif
        const double dt = (double)g * (1.0 / kBoysStepInv) - T;
        const double* row = tab + g * kBoysCols + k;
        double s = row[6];
        s = row[5] + s * dt * (1.0 / 6.0);
        s = row[4] + s * dt * (1.0 / 5.0);
        s = row[3] + s * dt * (1.0 / 4.0);
        s = row[2] + s * dt * (1.0 / 3.0);
        s = row[1] + s * dt * (1.0 / 2.0);
        return row[0] + s * dt;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = rsqrt(T);
#else
    const double r = 1.0 / std::sqrt(T);  // host evaluation of the same program (tests)
#endif
    double f = 0.88622692545275801365 * r;  // sqrt(pi)/2
    const double eT = exp(-T), h = 0.5 * r * r;
    for (int m = 0; m < k; ++m) f = ((2 * m + 1) * f - eT) * h;
    return f;
}

// Real solid harmonics of x^L-normalised Cartesians (the coefficients of ylm_coef): for l <= 2
// and spherical index s, up to three (Cartesian index, weight) terms.
struct SphTerms {
    int n;
    int c[3];
    double w[3];
};

__host__ __device__ inline SphTerms sph_terms(int l, int s) {
    const double r3 = 1.7320508075688772935;
    if (l == 1) return SphTerms{1, {s == 0 ? 1 : (s == 1 ? 2 : 0), 0, 0}, {1.0, 0.0, 0.0}};
    if (l == 2) {
        switch (s) {
            case 0: return SphTerms{1, {1, 0, 0}, {r3, 0.0, 0.0}};
            case 1: return SphTerms{1, {4, 0, 0}, {r3, 0.0, 0.0}};
            case 2: return SphTerms{3, {5, 0, 3}, {1.0, -0.5, -0.5}};
            case 3: return SphTerms{1, {2, 0, 0}, {r3, 0.0, 0.0}};
            default: return SphTerms{2, {0, 3, 0}, {0.5 * r3, -0.5 * r3, 0.0}};
        }
    }
    return SphTerms{1, {0, 0, 0}, {1.0, 0.0, 0.0}};
}

// One index (position pos, angular momentum l) of a block with dims in[0..3] -> spherical.
template <typename R>
HFG_INLINE void coop_sph_pass(const R* in, R* out, const int (&din)[4], int pos, int l) {
    int dout[4] = {din[0], din[1], din[2], din[3]};
    dout[pos] = 2 * l + 1;
    const int tot = dout[0] * dout[1] * dout[2] * dout[3];
    for (int o = threadIdx.x; o < tot; o += blockDim.x) {
        int idx[4];
        idx[3] = o % dout[3];
        idx[2] = (o / dout[3]) % dout[2];
        idx[1] = (o / (dout[3] * dout[2])) % dout[1];
        idx[0] = o / (dout[3] * dout[2] * dout[1]);
        const SphTerms t = sph_terms(l, idx[pos]);
        R acc = 0;
        for (int k = 0; k < t.n; ++k) {
            idx[pos] = t.c[k];
            acc += (R)t.w[k] * in[((idx[0] * din[1] + idx[1]) * din[2] + idx[2]) * din[3] + idx[3]];
        }
        out[o] = acc;
    }
}

template <typename R, typename ACC>
__global__ __launch_bounds__(kCoopBlock) void jk_coop_kernel(JkArgs a, CoopArgs p) {
    extern __shared__ double smem[];
    ACC* tiles = reinterpret_cast<ACC*>(smem);
    R* work = reinterpret_cast<R*>(smem + kRegions * kBlockFuncs * kBlockFuncs);
    R* H = work;
    R* bra = work + p.off_bra;
    R* Ic = work + p.off_I;
    R* Dblk = work + p.off_D;
    __shared__ int rowpre[kMaxRows + 1];
    __shared__ int next_item;
    const int NA = p.NA, NB = p.NB, NC = p.NC, ND = p.ND;
    const int na = p.na, nb = p.nb, nc = p.nc, nd = p.nd;
    const int NF = p.NF;
    const int lane = threadIdx.x, nl = blockDim.x;
    for (;;) {
        __syncthreads();
        if (lane == 0) next_item = atomicAdd(a.queue, 1);
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
        for (int k = lane; k < ntile; k += nl) tiles[k] = (ACC)0;
        const int nrows = w.r1 - w.r0;
        const int bpo = a.bp_off[w.P], kpo = a.bp_off[w.Q], kn = a.bp_cnt[w.Q];
        const bool same = (w.P == w.Q);
        for (int r = lane; r < nrows; r += nl) {
            const int i = w.r0 + r;
            int c = ket_count(a.psmean + kpo, kn, a.psmean[bpo + i], a.tol);
            if (same) c = min(c, i + 1);
            rowpre[r + 1] = max(0, min(c, w.c1) - w.c0);
        }
        __syncthreads();
        if (lane == 0) {
            rowpre[0] = 0;
            for (int r = 0; r < nrows; ++r) rowpre[r + 1] += rowpre[r];
        }
        __syncthreads();
        const int total = rowpre[nrows];
        for (int t = 0; t < total; ++t) {  // one quartet at a time, every lane on it
            int i, j;
            locate(t, rowpre, nrows, w, i, j);
            const PairDev bp = a.pairs[bpo + i];
            const PairDev kp = a.pairs[kpo + j];
            if (a.dtol > 0.0 && density_bound(bp, kp, dm, md.nshell) < a.dtol) continue;  // uniform
            double s = 1.0;
            if (bp.sa == bp.sb) s *= 0.5;
            if (kp.sa == kp.sb) s *= 0.5;
            if (same && i == j) s *= 0.5;
            R acc[kCoopMaxAcc];
#pragma unroll
            for (int q = 0; q < kCoopMaxAcc; ++q) acc[q] = (R)0;
            for (int ip = 0; ip < bp.np; ++ip) {
                const int ia = bp.off + ip * bp.pstride;
                const double4 a0 = a.pp0[ia];
                const double4 a1 = a.pp1[ia];
                const double o2p = a.pp2[ia].x;
                const double pp = a0.x;
                for (int jp = 0; jp < kp.np; ++jp) {
                    const int ib = kp.off + jp * kp.pstride;
                    const double4 b0 = a.pp0[ib];
                    const double4 b1 = a.pp1[ib];
                    const double o2q = a.pp2[ib].x;
                    const double q = b0.x;
                    const double rs = rsqrt(pp + q);
                    const double pfac = a1.w * b1.w * rs;
                    if (fabs(pfac) < a.prim_cut) continue;  // uniform
                    const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
                    const double ipq = rs * rs;
                    const double T = pp * q * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
                    const double rp = q * ipq, rq = pp * ipq;
                    if (lane <= p.LT) H[lane] = (R)(pfac * boys_order(lane, T, a.boys));
                    const R PA[3] = {(R)a1.x, (R)a1.y, (R)a1.z};
                    const R QC[3] = {(R)b1.x, (R)b1.y, (R)b1.z};
                    const R WP[3] = {(R)(-rp * PQx), (R)(-rp * PQy), (R)(-rp * PQz)};
                    const R WQ[3] = {(R)(rq * PQx), (R)(rq * PQy), (R)(rq * PQz)};
                    const R ro2p = (R)o2p, rrp = (R)rp, ro2q = (R)o2q, rrq = (R)rq, ro2pq = (R)(0.5 * ipq);
                    for (int l = 0; l < p.nvlev; ++l) {
                        __syncthreads();
                        for (int k = p.vlev[l] + lane; k < p.vlev[l + 1]; k += nl) {
                            const int4 op = p.vops[k];
                            const int tgt = op.x & 0xffff, d = (op.x >> 16) & 3, ket = (op.x >> 18) & 1;
                            const R n1 = (R)((op.x >> 19) & 15), ne = (R)((op.x >> 23) & 15);
                            R v;
                            if (!ket) {
                                v = PA[d] * H[op.y] + WP[d] * H[op.y + 1];
                                if (op.z >= 0) v += n1 * ro2p * (H[op.z] - rrp * H[op.z + 1]);
                            } else {
                                v = QC[d] * H[op.y] + WQ[d] * H[op.y + 1];
                                if (op.z >= 0) v += n1 * ro2q * (H[op.z] - rrq * H[op.z + 1]);
                                if (op.w >= 0) v += ne * ro2pq * H[op.w + 1];
                            }
                            H[tgt] = v;
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int q2 = 0; q2 < kCoopMaxAcc; ++q2) {
                        const int k = lane + q2 * kCoopBlock;
                        if (k < p.nacc) acc[q2] += H[p.acc_src[k]];
                    }
                    __syncthreads();  // the next primitive's Boys values overwrite H
                }
            }
            // bra HRR tables (one per ket function f) from the contracted values
#pragma unroll
            for (int q2 = 0; q2 < kCoopMaxAcc; ++q2) {
                const int k = lane + q2 * kCoopBlock;
                if (k < p.nacc) H[(k % NF) * p.hbsz + p.hb_init[k / NF]] = acc[q2];
            }
            const R AB[3] = {(R)bp.AB[0], (R)bp.AB[1], (R)bp.AB[2]};
            const R CD[3] = {(R)kp.AB[0], (R)kp.AB[1], (R)kp.AB[2]};
            for (int l = 0; l < p.nhblev; ++l) {
                __syncthreads();
                const int n0 = p.hblev[l], cnt = p.hblev[l + 1] - n0;
                for (int k = lane; k < cnt * NF; k += nl) {
                    const int f = k / cnt;
                    const int4 op = p.hbops[n0 + k - f * cnt];
                    R* h = H + f * p.hbsz;
                    h[op.x & 0xffff] = h[op.y] + AB[(op.x >> 16) & 3] * h[op.z];
                }
            }
            __syncthreads();
            for (int k = lane; k < p.nhb_out * NF; k += nl) {
                const int f = k / p.nhb_out;
                const int2 o = p.hb_out[k - f * p.nhb_out];
                bra[o.x * NF + f] = H[f * p.hbsz + o.y];
            }
            __syncthreads();
            // ket HRR tables (one per bra function pair), hk_chunk pairs at a time
            const int nab = na * nb, ncd = nc * nd;
            for (int ab0 = 0; ab0 < nab; ab0 += p.hk_chunk) {
                const int nch = min(p.hk_chunk, nab - ab0);
                for (int k = lane; k < nch * NF; k += nl) {
                    const int ab = k / NF, fi = k - ab * NF;
                    H[ab * p.hksz + p.hk_init[fi]] = bra[ab0 * NF + k];
                }
                for (int l = 0; l < p.nhklev; ++l) {
                    __syncthreads();
                    const int n0 = p.hklev[l], cnt = p.hklev[l + 1] - n0;
                    for (int k = lane; k < cnt * nch; k += nl) {
                        const int ab = k / cnt;
                        const int4 op = p.hkops[n0 + k - ab * cnt];
                        R* h = H + ab * p.hksz;
                        h[op.x & 0xffff] = h[op.y] + CD[(op.x >> 16) & 3] * h[op.z];
                    }
                }
                __syncthreads();
                for (int k = lane; k < p.nhk_out * nch; k += nl) {
                    const int ab = k / p.nhk_out;
                    const int2 o = p.hk_out[k - ab * p.nhk_out];
                    Ic[(ab0 + ab) * ncd + o.x] = H[ab * p.hksz + o.y];
                }
                __syncthreads();
            }
            // hfscf's element screen (components whose Schwarz bound is below the threshold) and,
            // for a Cartesian basis, the diagonal (pq|pq) taken from the screen matrix
            if (a.elem_tol > 0.0 && bp.sq_min * kp.sq_min < a.elem_tol) {
                const double* sb = a.psq + bp.sq_off;
                const double* sk = a.psq + kp.sq_off;
                for (int k = lane; k < nab * ncd; k += nl) {
                    const int ab = k / ncd;
                    Ic[k] *= (sb[ab] * sk[k - ab * ncd] >= a.elem_tol) ? (R)1 : (R)0;
                }
            }
            if (!p.pure && a.elem_tol > 0.0 && p.la == p.lc && p.lb == p.ld && bp.sa == kp.sa && bp.sb == kp.sb) {
                const double* s2 = a.psq + bp.sq_off;
                __syncthreads();
                for (int ab = lane; ab < nab; ab += nl) Ic[ab * nab + ab] = (R)(s2[ab] * s2[ab]);
            }
            // density blocks of the quartet
            const int ao_a = bp.ao_a, ao_b = bp.ao_b, ao_c = kp.ao_a, ao_d = kp.ao_b;
            R* dab = Dblk;
            R* dcd = dab + NA * NB;
            R* dac = dcd + NC * ND;
            R* dad = dac + NA * NC;
            R* dbc = dad + NA * ND;
            R* dbd = dbc + NB * NC;
            for (int k = lane; k < NA * NB; k += nl) dab[k] = (R)Dm[(size_t)(ao_a + k / NB) * md.n + ao_b + k % NB];
            for (int k = lane; k < NC * ND; k += nl) dcd[k] = (R)Dm[(size_t)(ao_c + k / ND) * md.n + ao_d + k % ND];
            for (int k = lane; k < NA * NC; k += nl) dac[k] = (R)Dm[(size_t)(ao_a + k / NC) * md.n + ao_c + k % NC];
            for (int k = lane; k < NA * ND; k += nl) dad[k] = (R)Dm[(size_t)(ao_a + k / ND) * md.n + ao_d + k % ND];
            for (int k = lane; k < NB * NC; k += nl) dbc[k] = (R)Dm[(size_t)(ao_b + k / NC) * md.n + ao_c + k % NC];
            for (int k = lane; k < NB * ND; k += nl) dbd[k] = (R)Dm[(size_t)(ao_b + k / ND) * md.n + ao_d + k % ND];
            // spherical transform, one index at a time (temporaries in H; the result in the
            // Cartesian region, free after the first pass)
            const R* I = Ic;
            if (p.pure) {
                int dims[4] = {na, nb, nc, nd};
                const int L4[4] = {p.la, p.lb, p.lc, p.ld};
                R* bufs[2] = {H, H + NA * nb * nc * nd};
                const R* src = Ic;
                int used = 0;
                for (int pos = 0; pos < 4; ++pos) {
                    if (L4[pos] == 0) continue;
                    ++used;
                    __syncthreads();
                    R* dst = bufs[(used - 1) & 1];
                    coop_sph_pass<R>(src, dst, dims, pos, L4[pos]);
                    dims[pos] = 2 * L4[pos] + 1;
                    src = dst;
                }
                I = src;
            }
            __syncthreads();
            if (a.eri_out) {
                // export: every canonical element (i >= j, k >= l, ij >= kl; hfscf's index_ijkl,
                // hfscf: src/math/hfscf_math.hpp:90) of the quartet, once. Each unordered pair of
                // shell pairs is one quartet of the lists, so only a shell paired with itself (i < j
                // within it) or a shell pair paired with itself (ij < kl) repeats an element.
                const bool same_ab = bp.sa == bp.sb, same_cd = kp.sa == kp.sb, same_q = same && i == j;
                auto tri = [](long long x, long long y) { return x >= y ? x * (x + 1) / 2 + y : y * (y + 1) / 2 + x; };
                for (int o = lane; o < NA * NB * NC * ND; o += nl) {
                    const int u = o % ND, r = (o / ND) % NC, qb = (o / (ND * NC)) % NB, pa = o / (ND * NC * NB);
                    const int fi = bp.ao_a + pa, fj = bp.ao_b + qb, fk = kp.ao_a + r, fl = kp.ao_b + u;
                    if ((same_ab && fi < fj) || (same_cd && fk < fl)) continue;
                    const long long ij = tri(fi, fj), kl = tri(fk, fl);
                    if (same_q && ij < kl) continue;
                    a.eri_out[tri(ij, kl)] = (double)I[o];
                }
                __syncthreads();  // the next quartet reuses every LDS region
                continue;
            }
            // digestion: J tiles 2 s D I, K tiles s D I (J = T + T^T, K = T + T^T as in jk_kernel)
            const R sj = (R)(2.0 * s), sk = (R)s;
            const int la0 = bp.slot_a * NA, lb0 = bp.slot_b * NB, lc0 = kp.slot_a * NC, ld0 = kp.slot_b * ND;
            const int n1 = NA * NB, n2 = n1 + NC * ND, n3 = n2 + NA * NC, n4 = n3 + NA * ND, n5 = n4 + NB * NC,
                      n6 = n5 + NB * ND;
            for (int o = lane; o < n6; o += nl) {
                R v = 0;
                if (o < n1) {  // J_ab
                    const int pa = o / NB, qb = o % NB;
                    for (int r = 0; r < NC; ++r)
                        for (int u = 0; u < ND; ++u) v += dcd[r * ND + u] * I[((pa * NB + qb) * NC + r) * ND + u];
                    tJab[(la0 + pa) * fB + lb0 + qb] += (ACC)(sj * v);
                } else if (o < n2) {  // J_cd
                    const int r = (o - n1) / ND, u = (o - n1) % ND;
                    for (int pa = 0; pa < NA; ++pa)
                        for (int qb = 0; qb < NB; ++qb) v += dab[pa * NB + qb] * I[((pa * NB + qb) * NC + r) * ND + u];
                    tJcd[(lc0 + r) * fD + ld0 + u] += (ACC)(sj * v);
                } else if (o < n3) {  // K_ac
                    const int pa = (o - n2) / NC, r = (o - n2) % NC;
                    for (int qb = 0; qb < NB; ++qb)
                        for (int u = 0; u < ND; ++u) v += dbd[qb * ND + u] * I[((pa * NB + qb) * NC + r) * ND + u];
                    tKac[(la0 + pa) * fC + lc0 + r] += (ACC)(sk * v);
                } else if (o < n4) {  // K_ad
                    const int pa = (o - n3) / ND, u = (o - n3) % ND;
                    for (int qb = 0; qb < NB; ++qb)
                        for (int r = 0; r < NC; ++r) v += dbc[qb * NC + r] * I[((pa * NB + qb) * NC + r) * ND + u];
                    tKad[(la0 + pa) * fD + ld0 + u] += (ACC)(sk * v);
                } else if (o < n5) {  // K_bc
                    const int qb = (o - n4) / NC, r = (o - n4) % NC;
                    for (int pa = 0; pa < NA; ++pa)
                        for (int u = 0; u < ND; ++u) v += dad[pa * ND + u] * I[((pa * NB + qb) * NC + r) * ND + u];
                    tKbc[(lb0 + qb) * fC + lc0 + r] += (ACC)(sk * v);
                } else {  // K_bd
                    const int qb = (o - n5) / ND, u = (o - n5) % ND;
                    for (int pa = 0; pa < NA; ++pa)
                        for (int r = 0; r < NC; ++r) v += dac[pa * NC + r] * I[((pa * NB + qb) * NC + r) * ND + u];
                    tKbd[(lb0 + qb) * fD + ld0 + u] += (ACC)(sk * v);
                }
            }
            __syncthreads();  // the next quartet reuses every LDS region
        }
        __syncthreads();
        if (a.eri_out) continue;  // an export has no tiles (the item may have no slab)
        double* out = a.slab + w.slab;
        for (int k = lane; k < ntile; k += nl) out[k] = (double)tiles[k];
    }
}

// Classes that run the cooperative kernel (J/K).
HFG_HD bool jk_uses_coop(int la, int lb, int lc, int ld) { return la + lb + lc + ld >= 7; }

// Host: the program and LDS layout of one class (arrays into `prog`, uploaded by the caller).
struct CoopHostProgram {
    CoopArgs args;                 // pointers relative to the start of the device buffer
    std::vector<int4> words;       // every array of the program, 16-byte aligned
    size_t lds_bytes(bool fp32) const;
};
CoopHostProgram make_coop_program(int la, int lb, int lc, int ld, bool pure);
// Rebases the program's array offsets onto the uploaded buffer.
CoopArgs coop_args_on_device(const CoopHostProgram& h, const int4* dev);

hipError_t launch_jk_coop(const JkArgs& a, const CoopArgs& p, size_t lds_bytes, int mixed, int nitems,
                          LaunchShape sh, hipStream_t s);
// The largest dynamic LDS the cooperative J/K and gradient launches request (set once per process
// before any launch; the attribute is per kernel).
hipError_t coop_set_lds_limit(size_t jk_bytes, size_t grad_bytes);

}  // namespace hfg
