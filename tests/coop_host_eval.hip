// Host evaluation of the cooperative J/K kernel's programs for one shell quartet (test helper,
// built by tests/test_coop_programs.py): the VRR, contraction, HRR and spherical steps of
// jk_coop_kernel run sequentially level by level on the CPU from the same programs.
#include "../hfscf_amd/csrc/gpu/coop.hip"
#include "../hfscf_amd/csrc/gpu/boys_table.hip"
#include <cmath>

using namespace hfg;

static std::vector<double> g_tab;

extern "C" int coop_eval(int la, int lb, int lc, int ld, int pure, const double* cen /*4x3*/, const int* np /*4*/,
                          const double* alpha, const double* coef /*concatenated, normalised*/, double* out) {
    if (g_tab.empty()) g_tab = boys_table();
    CoopHostProgram hp = make_coop_program(la, lb, lc, ld, pure != 0);
    CoopArgs p = coop_args_on_device(hp, hp.words.data());
    const double* A = cen; const double* B = cen + 3; const double* C = cen + 6; const double* D = cen + 9;
    int o[4] = {0, np[0], np[0] + np[1], np[0] + np[1] + np[2]};
    std::vector<double> H((size_t)p.lds_r + 16, 0.0), acc(p.nacc, 0.0);
    double AB[3], CD[3];
    for (int d = 0; d < 3; ++d) { AB[d] = A[d] - B[d]; CD[d] = C[d] - D[d]; }
    const double k2 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    for (int i = 0; i < np[0]; ++i) for (int j = 0; j < np[1]; ++j) {
        double a = alpha[o[0] + i], b = alpha[o[1] + j], pp = a + b;
        double P[3], PA[3], ab2 = 0;
        for (int d = 0; d < 3; ++d) { P[d] = (a * A[d] + b * B[d]) / pp; PA[d] = P[d] - A[d]; ab2 += AB[d] * AB[d]; }
        double Kab = std::exp(-a * b / pp * ab2) * coef[o[0] + i] * coef[o[1] + j] * k2 / pp;
        for (int k = 0; k < np[2]; ++k) for (int l = 0; l < np[3]; ++l) {
            double c = alpha[o[2] + k], e = alpha[o[3] + l], q = c + e;
            double Q[3], QC[3], cd2 = 0;
            for (int d = 0; d < 3; ++d) { Q[d] = (c * C[d] + e * D[d]) / q; QC[d] = Q[d] - C[d]; cd2 += CD[d] * CD[d]; }
            double Kcd = std::exp(-c * e / q * cd2) * coef[o[2] + k] * coef[o[3] + l] * k2 / q;
            double rs = 1.0 / std::sqrt(pp + q), pfac = Kab * Kcd * rs;
            double PQ[3], r2 = 0; for (int d = 0; d < 3; ++d) { PQ[d] = P[d] - Q[d]; r2 += PQ[d] * PQ[d]; }
            double ipq = rs * rs, T = pp * q * ipq * r2, rp = q * ipq, rq = pp * ipq;
            double WP[3], WQ[3]; for (int d = 0; d < 3; ++d) { WP[d] = -rp * PQ[d]; WQ[d] = rq * PQ[d]; }
            double o2p = 0.5 / pp, o2q = 0.5 / q, o2pq = 0.5 * ipq;
            for (int m = 0; m <= p.LT; ++m) H[m] = pfac * boys_order(m, T, g_tab.data());
            for (int lv = 0; lv < p.nvlev; ++lv)
                for (int kk = p.vlev[lv]; kk < p.vlev[lv + 1]; ++kk) {
                    int4 op = p.vops[kk];
                    int tgt = op.x & 0xffff, d = (op.x >> 16) & 3, ket = (op.x >> 18) & 1;
                    double n1 = (op.x >> 19) & 15, ne = (op.x >> 23) & 15, v;
                    if (!ket) { v = PA[d] * H[op.y] + WP[d] * H[op.y + 1]; if (op.z >= 0) v += n1 * o2p * (H[op.z] - rp * H[op.z + 1]); }
                    else { v = QC[d] * H[op.y] + WQ[d] * H[op.y + 1]; if (op.z >= 0) v += n1 * o2q * (H[op.z] - rq * H[op.z + 1]);
                           if (op.w >= 0) v += ne * o2pq * H[op.w + 1]; }
                    H[tgt] = v;
                }
            for (int kk = 0; kk < p.nacc; ++kk) acc[kk] += H[p.acc_src[kk]];
        }
    }
    const int NF = p.NF, na = p.na, nb = p.nb, nc = p.nc, nd = p.nd, nab = na * nb, ncd = nc * nd;
    double* bra = H.data() + p.off_bra; double* Ic = H.data() + p.off_I; double* Hh = H.data();
    for (int kk = 0; kk < p.nacc; ++kk) Hh[(kk % NF) * p.hbsz + p.hb_init[kk / NF]] = acc[kk];
    for (int lv = 0; lv < p.nhblev; ++lv) {
        int n0 = p.hblev[lv], cnt = p.hblev[lv + 1] - n0;
        for (int kk = 0; kk < cnt * NF; ++kk) { int f = kk / cnt; int4 op = p.hbops[n0 + kk - f * cnt]; double* h = Hh + f * p.hbsz;
            h[op.x & 0xffff] = h[op.y] + AB[(op.x >> 16) & 3] * h[op.z]; }
    }
    for (int kk = 0; kk < p.nhb_out * NF; ++kk) { int f = kk / p.nhb_out; int2 oo = p.hb_out[kk - f * p.nhb_out]; bra[oo.x * NF + f] = Hh[f * p.hbsz + oo.y]; }
    for (int ab0 = 0; ab0 < nab; ab0 += p.hk_chunk) {
        int nch = std::min(p.hk_chunk, nab - ab0);
        for (int kk = 0; kk < nch * NF; ++kk) { int ab = kk / NF, fi = kk - ab * NF; Hh[ab * p.hksz + p.hk_init[fi]] = bra[ab0 * NF + kk]; }
        for (int lv = 0; lv < p.nhklev; ++lv) {
            int n0 = p.hklev[lv], cnt = p.hklev[lv + 1] - n0;
            for (int kk = 0; kk < cnt * nch; ++kk) { int ab = kk / cnt; int4 op = p.hkops[n0 + kk - ab * cnt]; double* h = Hh + ab * p.hksz;
                h[op.x & 0xffff] = h[op.y] + CD[(op.x >> 16) & 3] * h[op.z]; }
        }
        for (int kk = 0; kk < p.nhk_out * nch; ++kk) { int ab = kk / p.nhk_out; int2 oo = p.hk_out[kk - ab * p.nhk_out]; Ic[(ab0 + ab) * ncd + oo.x] = Hh[ab * p.hksz + oo.y]; }
    }
    const double* I = Ic;
    if (p.pure) {
        int dims[4] = {na, nb, nc, nd}; const int L4[4] = {p.la, p.lb, p.lc, p.ld};
        double* bufs[2] = {Hh, Hh + p.NA * nb * nc * nd}; const double* src = Ic; int used = 0;
        for (int pos = 0; pos < 4; ++pos) {
            if (L4[pos] == 0) continue; ++used; double* dst = bufs[(used - 1) & 1];
            int dout[4] = {dims[0], dims[1], dims[2], dims[3]}; dout[pos] = 2 * L4[pos] + 1;
            int tot = dout[0] * dout[1] * dout[2] * dout[3];
            for (int oo = 0; oo < tot; ++oo) {
                int idx[4]; idx[3] = oo % dout[3]; idx[2] = (oo / dout[3]) % dout[2]; idx[1] = (oo / (dout[3] * dout[2])) % dout[1]; idx[0] = oo / (dout[3] * dout[2] * dout[1]);
                SphTerms t = sph_terms(L4[pos], idx[pos]); double s = 0;
                for (int k = 0; k < t.n; ++k) { idx[pos] = t.c[k]; s += t.w[k] * src[((idx[0] * dims[1] + idx[1]) * dims[2] + idx[2]) * dims[3] + idx[3]]; }
                dst[oo] = s;
            }
            dims[pos] = 2 * L4[pos] + 1; src = dst;
        }
        I = src;
    }
    const int n = p.NA * p.NB * p.NC * p.ND;
    for (int kk = 0; kk < n; ++kk) out[kk] = I[kk];
    return (int)hp.lds_bytes(false);
}

// The cooperative gradient program for one shell quartet contracted with a Cartesian density
// block G (na nb nc nd): out[3 * centre + x] = sum G d(ab|cd)/dX_x for centres A, B, C, run
// sequentially from the program (VRR at L + 1, shifted-class HRR, contraction tables, weights).
extern "C" int coop_grad_eval(int la, int lb, int lc, int ld, const double* cen, const int* np, const double* alpha,
                              const double* coef, const double* G, double* out) {
    if (g_tab.empty()) g_tab = boys_table();
    CoopGradHostProgram hp = make_coop_grad_program(la, lb, lc, ld, false, 160 * 1024 - kCoopGradStaticLds);
    CoopGradArgs p = coop_grad_args_on_device(hp, hp.words.data());
    const double* A = cen; const double* B = cen + 3; const double* C = cen + 6; const double* D = cen + 9;
    int o[4] = {0, np[0], np[0] + np[1], np[0] + np[1] + np[2]};
    std::vector<double> L((size_t)p.lds_r + 16, 0.0);
    double* V = L.data() + p.off_V; double* Hb = L.data() + p.off_Hb; double* bra = L.data() + p.off_bra; double* Hk = L.data() + p.off_Hk;
    double AB[3], CD[3];
    for (int d = 0; d < 3; ++d) { AB[d] = A[d] - B[d]; CD[d] = C[d] - D[d]; }
    for (int k = 0; k < 9; ++k) out[k] = 0;
    const double k2 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    for (int i = 0; i < np[0]; ++i) for (int j = 0; j < np[1]; ++j) {
        double a = alpha[o[0] + i], b = alpha[o[1] + j], pp = a + b;
        double P[3], PA[3], ab2 = 0;
        for (int d = 0; d < 3; ++d) { P[d] = (a * A[d] + b * B[d]) / pp; PA[d] = P[d] - A[d]; ab2 += AB[d] * AB[d]; }
        double Kab = std::exp(-a * b / pp * ab2) * coef[o[0] + i] * coef[o[1] + j] * k2 / pp;
        for (int k = 0; k < np[2]; ++k) for (int l = 0; l < np[3]; ++l) {
            double c = alpha[o[2] + k], e = alpha[o[3] + l], q = c + e;
            double Q[3], QC[3], cd2 = 0;
            for (int d = 0; d < 3; ++d) { Q[d] = (c * C[d] + e * D[d]) / q; QC[d] = Q[d] - C[d]; cd2 += CD[d] * CD[d]; }
            double Kcd = std::exp(-c * e / q * cd2) * coef[o[2] + k] * coef[o[3] + l] * k2 / q;
            double rs = 1.0 / std::sqrt(pp + q), pfac = Kab * Kcd * rs;
            double PQ[3], r2 = 0; for (int d = 0; d < 3; ++d) { PQ[d] = P[d] - Q[d]; r2 += PQ[d] * PQ[d]; }
            double ipq = rs * rs, T = pp * q * ipq * r2, rp = q * ipq, rq = pp * ipq;
            double WP[3], WQ[3]; for (int d = 0; d < 3; ++d) { WP[d] = -rp * PQ[d]; WQ[d] = rq * PQ[d]; }
            double o2p = 0.5 / pp, o2q = 0.5 / q, o2pq = 0.5 * ipq;
            for (int m = 0; m <= p.LT + 1; ++m) V[m] = pfac * boys_order(m, T, g_tab.data());
            for (int lv = 0; lv < p.nvlev; ++lv)
                for (int kk = p.vlev[lv]; kk < p.vlev[lv + 1]; ++kk) {
                    int4 op = p.vops[kk];
                    int tgt = op.x & 0xffff, d = (op.x >> 16) & 3, ket = (op.x >> 18) & 1;
                    double n1 = (op.x >> 19) & 15, ne = (op.x >> 23) & 15, v;
                    if (!ket) { v = PA[d] * V[op.y] + WP[d] * V[op.y + 1]; if (op.z >= 0) v += n1 * o2p * (V[op.z] - rp * V[op.z + 1]); }
                    else { v = QC[d] * V[op.y] + WQ[d] * V[op.y + 1]; if (op.z >= 0) v += n1 * o2q * (V[op.z] - rq * V[op.z + 1]);
                           if (op.w >= 0) v += ne * o2pq * V[op.w + 1]; }
                    V[tgt] = v;
                }
            const double w3[3] = {2.0 * a, 2.0 * b, 2.0 * c};
            for (int si = 0; si < p.nshift; ++si) {
                const CoopShift& S = p.sh[si];
                const int NF = S.NF, nab = S.na * S.nb, ncd = S.nc * S.nd;
                for (int f0 = 0; f0 < NF; f0 += S.hb_chunk) {
                    const int nf = std::min(S.hb_chunk, NF - f0);
                    for (int kk = 0; kk < S.NE * nf; ++kk) { int ee = kk / nf, f = kk - ee * nf; Hb[f * S.hbsz + S.hb_init[ee]] = V[S.hb_src[ee * NF + f0 + f]]; }
                    for (int lv = 0; lv < S.nhblev; ++lv) { int n0 = S.hblev[lv], cnt = S.hblev[lv + 1] - n0;
                        for (int kk = 0; kk < cnt * nf; ++kk) { int f = kk / cnt; int4 op = S.hbops[n0 + kk - f * cnt]; double* h = Hb + f * S.hbsz;
                            h[op.x & 0xffff] = h[op.y] + AB[(op.x >> 16) & 3] * h[op.z]; } }
                    for (int kk = 0; kk < S.nhb_out * nf; ++kk) { int f = kk / S.nhb_out; int2 oo = S.hb_out[kk - f * S.nhb_out]; bra[oo.x * NF + f0 + f] = Hb[f * S.hbsz + oo.y]; }
                }
                double s3[3] = {0, 0, 0};
                for (int ab0 = 0; ab0 < nab; ab0 += S.hk_chunk) {
                    const int nch = std::min(S.hk_chunk, nab - ab0);
                    for (int kk = 0; kk < nch * NF; ++kk) { int ab = kk / NF, fi = kk - ab * NF; Hk[ab * S.hksz + S.hk_init[fi]] = bra[ab0 * NF + kk]; }
                    for (int lv = 0; lv < S.nhklev; ++lv) { int n0 = S.hklev[lv], cnt = S.hklev[lv + 1] - n0;
                        for (int kk = 0; kk < cnt * nch; ++kk) { int ab = kk / cnt; int4 op = S.hkops[n0 + kk - ab * cnt]; double* h = Hk + ab * S.hksz;
                            h[op.x & 0xffff] = h[op.y] + CD[(op.x >> 16) & 3] * h[op.z]; } }
                    for (int kk = 0; kk < S.nhk_out * nch; ++kk) { int ab = kk / S.nhk_out; int2 oo = S.hk_out[kk - ab * S.nhk_out];
                        double v = Hk[ab * S.hksz + oo.y]; const int* c3 = S.cg + 3 * ((ab0 + ab) * ncd + oo.x);
                        for (int x = 0; x < 3; ++x) if (c3[x] >= 0) s3[x] += (double)(c3[x] >> 20) * G[c3[x] & 0xfffff] * v; }
                }
                const double wt = S.shift > 0 ? w3[S.pos] : -1.0;
                for (int x = 0; x < 3; ++x) out[3 * S.pos + x] += wt * s3[x];
            }
        }
    }
    return (int)hp.lds_bytes(false);
}
