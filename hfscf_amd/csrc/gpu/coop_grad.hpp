// Wavefront-cooperative adjoint gradient kernel for the classes of highest angular momentum
// ((dd|dd), (dd|dp), (dp|dd)): the contraction of grad_kernel (grad_impl.hpp)
// with one workgroup per shell quartet and every step spread over its lanes, tables in LDS:
//   1. the two-particle density of the quartet in the output basis, back-transformed to
//      Cartesian components G (lanes over elements);
//   2. per primitive quartet: Boys values (one order per lane) and the VRR at L+1 from the same
//      kind of program as the J/K kernel;
//   3. per shifted class (a +- 1_i, b +- 1_i, c +- 1_i): the bra HRR per ket function and the
//      ket HRR per bra function pair from programs, each final integral I contracted with G as
//      it leaves the ket HRR (a table gives, per integral and direction, the G element and
//      factor: +1 shifts G[x - 1_i], -1 shifts (x_i + 1) G[x + 1_i]), so no shifted block is stored;
//      the lanes' sums, weighted by 2 alpha (+1) or -1 (-1), stay in registers;
//   4. one reduction per quartet into the per-atom sums in LDS (D by translational invariance).
#pragma once
#include <cmath>
#include <vector>

#include "coop.hpp"
#include "grad_impl.hpp"

namespace hfg {

// Cartesian component c of l as spherical terms: Y(c, s) != 0 (the transpose of sph_terms).
__host__ __device__ inline SphTerms cart_terms(int l, int c) {
    const double r3 = 1.7320508075688772935;
    if (l == 1) return SphTerms{1, {c == 0 ? 2 : (c == 1 ? 0 : 1), 0, 0}, {1.0, 0.0, 0.0}};
    if (l == 2) {
        switch (c) {
            case 0: return SphTerms{2, {2, 4, 0}, {-0.5, 0.5 * r3, 0.0}};
            case 1: return SphTerms{1, {0, 0, 0}, {r3, 0.0, 0.0}};
            case 2: return SphTerms{1, {3, 0, 0}, {r3, 0.0, 0.0}};
            case 3: return SphTerms{2, {2, 4, 0}, {-0.5, -0.5 * r3, 0.0}};
            case 4: return SphTerms{1, {1, 0, 0}, {r3, 0.0, 0.0}};
            default: return SphTerms{1, {2, 0, 0}, {1.0, 0.0, 0.0}};
        }
    }
    return SphTerms{1, {0, 0, 0}, {1.0, 0.0, 0.0}};
}

// One index (position pos, angular momentum l) of a block with dims din back to Cartesian.
template <typename R>
HFG_INLINE void coop_cart_pass(const R* in, R* out, const int (&din)[4], int pos, int l) {
    int dout[4] = {din[0], din[1], din[2], din[3]};
    dout[pos] = ncart(l);
    const int tot = dout[0] * dout[1] * dout[2] * dout[3];
    for (int o = threadIdx.x; o < tot; o += blockDim.x) {
        int idx[4];
        idx[3] = o % dout[3];
        idx[2] = (o / dout[3]) % dout[2];
        idx[1] = (o / (dout[3] * dout[2])) % dout[1];
        idx[0] = o / (dout[3] * dout[2] * dout[1]);
        const SphTerms t = cart_terms(l, idx[pos]);
        R acc = 0;
        for (int k = 0; k < t.n; ++k) {
            idx[pos] = t.c[k];
            acc += (R)t.w[k] * in[((idx[0] * din[1] + idx[1]) * din[2] + idx[2]) * din[3] + idx[3]];
        }
        out[o] = acc;
    }
}

HFG_INLINE double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <typename R>
__global__ __launch_bounds__(kCoopGradBlock) void grad_coop_kernel(GradArgs a, CoopGradArgs p) {
    extern __shared__ double smem[];
    R* G = reinterpret_cast<R*>(smem);
    R* V = G + p.off_V;
    R* Hb = G + p.off_Hb;
    R* bra = G + p.off_bra;
    R* Hk = G + p.off_Hk;
    __shared__ double gsum[kMaxAtoms * 3];
    __shared__ double red[9];
    __shared__ int rowpre[kMaxRows + 1];
    __shared__ int next_item;
    const int lane = threadIdx.x, nl = blockDim.x;
    const int NA = p.NA, NB = p.NB, NC = p.NC, ND = p.ND;
    for (;;) {
        __syncthreads();
        if (lane == 0) next_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = next_item;
        if (item >= a.nitems) break;
        const WorkItem w = a.items[item];
        const MolDev md = a.mols[w.mol];
        const double* __restrict__ Dm = a.D + md.moff;
        const int natom3 = md.natom * 3;
        for (int k = lane; k < natom3; k += nl) gsum[k] = 0.0;
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
        for (int t = 0; t < total; ++t) {
            int i, j;
            locate(t, rowpre, nrows, w, i, j);
            const PairDev bp = a.pairs[bpo + i];
            const PairDev kp = a.pairs[kpo + j];
            double deg = 8.0;
            if (bp.sa == bp.sb) deg *= 0.5;
            if (kp.sa == kp.sb) deg *= 0.5;
            if (same && i == j) deg *= 0.5;
            // two-particle density in the output basis (in V), then Cartesian components in G
            {
                const int oa = bp.ao_a, ob = bp.ao_b, oc = kp.ao_a, od = kp.ao_b, n = md.n;
                R* Gs = p.pure ? V : G;
                const int tot = NA * NB * NC * ND;
                for (int k = lane; k < tot; k += nl) {
                    const int s = k % ND, r = (k / ND) % NC, q = (k / (ND * NC)) % NB, pa = k / (ND * NC * NB);
                    const double dpq = Dm[(size_t)(oa + pa) * n + ob + q], drs = Dm[(size_t)(oc + r) * n + od + s];
                    const double dpr = Dm[(size_t)(oa + pa) * n + oc + r], dqs = Dm[(size_t)(ob + q) * n + od + s];
                    const double dps = Dm[(size_t)(oa + pa) * n + od + s], dqr = Dm[(size_t)(ob + q) * n + oc + r];
                    Gs[k] = (R)(deg * (2.0 * dpq * drs - 0.5 * (dpr * dqs + dps * dqr)));
                }
                if (p.pure) {
                    // passes alternate between V[0, nc4) (holding Gs first) and V[nc4, 2 nc4),
                    // nc4 = Cartesian block size >= every intermediate; the last pass writes G
                    int dims[4] = {NA, NB, NC, ND};
                    const int L4[4] = {p.la, p.lb, p.lc, p.ld};
                    const int nc4 = p.na * p.nb * p.nc * p.nd;
                    int last = -1;
                    for (int pos = 0; pos < 4; ++pos)
                        if (L4[pos] > 0) last = pos;
                    R* bufs[2] = {V + nc4, V};
                    const R* src = Gs;
                    int used = 0;
                    for (int pos = 0; pos < 4; ++pos) {
                        if (L4[pos] == 0) continue;
                        __syncthreads();
                        R* dst = pos == last ? G : bufs[used & 1];
                        ++used;
                        coop_cart_pass<R>(src, dst, dims, pos, L4[pos]);
                        dims[pos] = ncart(L4[pos]);
                        src = dst;
                    }
                }
            }
            // d/dA, d/dB, d/dC (x, y, z each): this lane's share
            double gA0 = 0, gA1 = 0, gA2 = 0, gB0 = 0, gB1 = 0, gB2 = 0, gC0 = 0, gC1 = 0, gC2 = 0;
            const R AB[3] = {(R)bp.AB[0], (R)bp.AB[1], (R)bp.AB[2]};
            const R CD[3] = {(R)kp.AB[0], (R)kp.AB[1], (R)kp.AB[2]};
            for (int ip = 0; ip < bp.np; ++ip) {
                const int ia = bp.off + ip * bp.pstride;
                const double4 a0 = a.pp0[ia];
                const double4 a1 = a.pp1[ia];
                const double4 a2 = a.pp2[ia];
                const double pp = a0.x;
                for (int jp = 0; jp < kp.np; ++jp) {
                    const int ib = kp.off + jp * kp.pstride;
                    const double4 b0 = a.pp0[ib];
                    const double4 b1 = a.pp1[ib];
                    const double4 b2 = a.pp2[ib];
                    const double q = b0.x;
                    const double rs = rsqrt(pp + q);
                    const double pfac = a1.w * b1.w * rs;
                    if (fabs(pfac) < a.prim_cut) continue;  // uniform
                    const double PQx = a0.y - b0.y, PQy = a0.z - b0.z, PQz = a0.w - b0.w;
                    const double ipq = rs * rs;
                    const double T = pp * q * ipq * (PQx * PQx + PQy * PQy + PQz * PQz);
                    const double rp = q * ipq, rq = pp * ipq;
                    __syncthreads();  // G complete; the previous primitive is done with V
                    if (lane <= p.LT + 1) V[lane] = (R)(pfac * boys_order(lane, T, a.boys));
                    const R PA[3] = {(R)a1.x, (R)a1.y, (R)a1.z};
                    const R QC[3] = {(R)b1.x, (R)b1.y, (R)b1.z};
                    const R WP[3] = {(R)(-rp * PQx), (R)(-rp * PQy), (R)(-rp * PQz)};
                    const R WQ[3] = {(R)(rq * PQx), (R)(rq * PQy), (R)(rq * PQz)};
                    const R ro2p = (R)a2.x, rrp = (R)rp, ro2q = (R)b2.x, rrq = (R)rq, ro2pq = (R)(0.5 * ipq);
                    for (int l = 0; l < p.nvlev; ++l) {
                        __syncthreads();
                        for (int k = p.vlev[l] + lane; k < p.vlev[l + 1]; k += nl) {
                            const int4 op = p.vops[k];
                            const int tgt = op.x & 0xffff, d = (op.x >> 16) & 3, ket = (op.x >> 18) & 1;
                            const R n1 = (R)((op.x >> 19) & 15), ne = (R)((op.x >> 23) & 15);
                            R v;
                            if (!ket) {
                                v = PA[d] * V[op.y] + WP[d] * V[op.y + 1];
                                if (op.z >= 0) v += n1 * ro2p * (V[op.z] - rrp * V[op.z + 1]);
                            } else {
                                v = QC[d] * V[op.y] + WQ[d] * V[op.y + 1];
                                if (op.z >= 0) v += n1 * ro2q * (V[op.z] - rrq * V[op.z + 1]);
                                if (op.w >= 0) v += ne * ro2pq * V[op.w + 1];
                            }
                            V[tgt] = v;
                        }
                    }
                    const double wa = 2.0 * a2.y, wb = 2.0 * a2.z, wc = 2.0 * b2.y;
                    for (int si = 0; si < p.nshift; ++si) {
                        const CoopShift& S = p.sh[si];
                        const int NF = S.NF, nab = S.na * S.nb, ncd = S.nc * S.nd;
                        // bra HRR, hb_chunk ket functions at a time
                        for (int f0 = 0; f0 < NF; f0 += S.hb_chunk) {
                            const int nf = min(S.hb_chunk, NF - f0);
                            __syncthreads();  // VRR complete / the previous chunk's tables read
                            for (int k = lane; k < S.NE * nf; k += nl) {
                                const int e = k / nf, f = k - e * nf;
                                Hb[f * S.hbsz + S.hb_init[e]] = V[S.hb_src[e * NF + f0 + f]];
                            }
                            for (int l = 0; l < S.nhblev; ++l) {
                                __syncthreads();
                                const int n0 = S.hblev[l], cnt = S.hblev[l + 1] - n0;
                                for (int k = lane; k < cnt * nf; k += nl) {
                                    const int f = k / cnt;
                                    const int4 op = S.hbops[n0 + k - f * cnt];
                                    R* h = Hb + f * S.hbsz;
                                    h[op.x & 0xffff] = h[op.y] + AB[(op.x >> 16) & 3] * h[op.z];
                                }
                            }
                            __syncthreads();
                            for (int k = lane; k < S.nhb_out * nf; k += nl) {
                                const int f = k / S.nhb_out;
                                const int2 o = S.hb_out[k - f * S.nhb_out];
                                bra[o.x * NF + f0 + f] = Hb[f * S.hbsz + o.y];
                            }
                        }
                        // ket HRR per bra function pair, each integral contracted with G at once
                        double s0 = 0, s1 = 0, s2 = 0;
                        for (int ab0 = 0; ab0 < nab; ab0 += S.hk_chunk) {
                            const int nch = min(S.hk_chunk, nab - ab0);
                            __syncthreads();
                            for (int k = lane; k < nch * NF; k += nl) {
                                const int ab = k / NF, fi = k - ab * NF;
                                Hk[ab * S.hksz + S.hk_init[fi]] = bra[ab0 * NF + k];
                            }
                            for (int l = 0; l < S.nhklev; ++l) {
                                __syncthreads();
                                const int n0 = S.hklev[l], cnt = S.hklev[l + 1] - n0;
                                for (int k = lane; k < cnt * nch; k += nl) {
                                    const int ab = k / cnt;
                                    const int4 op = S.hkops[n0 + k - ab * cnt];
                                    R* h = Hk + ab * S.hksz;
                                    h[op.x & 0xffff] = h[op.y] + CD[(op.x >> 16) & 3] * h[op.z];
                                }
                            }
                            __syncthreads();
                            for (int k = lane; k < S.nhk_out * nch; k += nl) {
                                const int ab = k / S.nhk_out;
                                const int2 o = S.hk_out[k - ab * S.nhk_out];
                                const double v = (double)Hk[ab * S.hksz + o.y];
                                const int* c3 = S.cg + 3 * ((ab0 + ab) * ncd + o.x);
                                const int g0 = c3[0], g1 = c3[1], g2 = c3[2];
                                if (g0 >= 0) s0 += (double)(g0 >> 20) * (double)G[g0 & 0xfffff] * v;
                                if (g1 >= 0) s1 += (double)(g1 >> 20) * (double)G[g1 & 0xfffff] * v;
                                if (g2 >= 0) s2 += (double)(g2 >> 20) * (double)G[g2 & 0xfffff] * v;
                            }
                        }
                        const double wt = S.shift > 0 ? (S.pos == 0 ? wa : (S.pos == 1 ? wb : wc)) : -1.0;
                        if (S.pos == 0) {
                            gA0 += wt * s0;
                            gA1 += wt * s1;
                            gA2 += wt * s2;
                        } else if (S.pos == 1) {
                            gB0 += wt * s0;
                            gB1 += wt * s1;
                            gB2 += wt * s2;
                        } else {
                            gC0 += wt * s0;
                            gC1 += wt * s1;
                            gC2 += wt * s2;
                        }
                    }
                }
            }
            // one reduction per quartet into the per-atom sums (d/dD = -(d/dA + d/dB + d/dC))
            __syncthreads();
            if (lane < 9) red[lane] = 0.0;
            __syncthreads();
            {
                const double part[9] = {gA0, gA1, gA2, gB0, gB1, gB2, gC0, gC1, gC2};
#pragma unroll
                for (int x = 0; x < 9; ++x) {
                    const double v = wave_sum(part[x]);
                    if ((lane & 63) == 0) atomicAdd(&red[x], v);
                }
            }
            __syncthreads();
            if (lane == 0) {
                const int* sat = a.shell_atom + md.shell0;
                const int atA = sat[bp.sa], atB = sat[bp.sb], atC = sat[kp.sa], atD = sat[kp.sb];
                for (int d = 0; d < 3; ++d) {
                    const double xa = red[d], xb = red[3 + d], xc = red[6 + d];
                    gsum[atA * 3 + d] += xa;
                    gsum[atB * 3 + d] += xb;
                    gsum[atC * 3 + d] += xc;
                    gsum[atD * 3 + d] -= xa + xb + xc;
                }
            }
        }
        __syncthreads();
        double* out = a.slab + w.gslab;
        for (int k = lane; k < natom3; k += nl) out[k] = gsum[k];
    }
}

// Classes that run the cooperative gradient kernel.
HFG_HD bool grad_uses_coop(int la, int lb, int lc, int ld) { return la + lb + lc + ld >= 7; }

struct CoopGradHostProgram {
    CoopGradArgs args;
    std::vector<int4> words;
    size_t lds_bytes(bool fp32) const { return (fp32 ? sizeof(float) : sizeof(double)) * (size_t)args.lds_r; }
};
// The program of a gradient class whose FP64 LDS stays within lds_limit bytes (the chunks of the
// HRR tables are chosen for it).
CoopGradHostProgram make_coop_grad_program(int la, int lb, int lc, int ld, bool pure, size_t lds_limit);
CoopGradArgs coop_grad_args_on_device(const CoopGradHostProgram& h, const int4* dev);
hipError_t launch_grad_coop(const GradArgs& a, const CoopGradArgs& p, size_t lds_bytes, int mixed, int nitems,
                            LaunchShape sh, hipStream_t s);

}  // namespace hfg
