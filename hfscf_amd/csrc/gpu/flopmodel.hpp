// Algorithmic FLOP model v1 for one contracted shell quartet (la lb|lc ld) with K primitive
// quartets; the same definition as hfscf_amd/flopmodel.py:
//   F = K [30 (setup) + (30 + 3L) (Boys + downward) + VRR + 2 N_con] + HRR + SPH + 12 na nb nc nd
//   VRR: 2 FLOP per Obara-Saika term building (e0|f0)^(m), |e| <= la+lb, |f| <= lc+ld,
//        m <= L - |e| - |f|; bra targets have 2 + 2[e_d - 1 > 0] terms, ket targets
//        2 + 2[f_d - 1 > 0] + [e_d > 0], d = first nonzero direction of the target.
//   N_con = sum_{A=la}^{la+lb} ncart(A) * sum_{C=lc}^{lc+ld} ncart(C)
//   HRR: 2 FLOP per bra-HRR and ket-HRR target; SPH: 2 nnz(l) x (other three current dims)
//   per spherical index with l >= 2 (nnz(d) = 8); digestion 12 FLOP per function quartet.
// The gradient model uses VRR at (la+lb+1, lc+ld+1, L+1), N_con over the shifted ranges,
// 3x HRR/SPH and 26 FLOP per function quartet.
#pragma once
#include <vector>

namespace hfg {

inline int fm_ncart(int L) { return (L + 1) * (L + 2) / 2; }

struct FlopParts {
    double per_prim;     // FLOPs per primitive quartet
    double per_quartet;  // FLOPs per contracted quartet
};

inline double vrr_flops(int LAB, int LCD, int LT) {
    double f = 0;
    // enumerate components (lx, ly, lz)
    for (int le = 0; le <= LAB; ++le)
        for (int i = 0; i <= le; ++i)
            for (int j = 0; j <= i; ++j) {
                const int e[3] = {le - i, i - j, j};
                const int d = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                if (le >= 1) {  // bra target
                    const int terms = 2 + 2 * (e[d] - 1 > 0);
                    f += 2.0 * terms * (LT - le + 1);
                }
                for (int lf = 1; lf <= LCD; ++lf)
                    for (int a = 0; a <= lf; ++a)
                        for (int b = 0; b <= a; ++b) {
                            const int g[3] = {lf - a, a - b, b};
                            const int dd = g[0] > 0 ? 0 : (g[1] > 0 ? 1 : 2);
                            const int mcount = LT - le - lf + 1;
                            if (mcount <= 0) continue;
                            const int terms = 2 + 2 * (g[dd] - 1 > 0) + (e[dd] > 0);
                            f += 2.0 * terms * mcount;
                        }
            }
    return f;
}

inline FlopParts model_flops(int la, int lb, int lc, int ld, bool pure, bool gradient) {
    const int L = la + lb + lc + ld;
    const int sh = gradient ? 1 : 0;
    const int LAB = la + lb + sh, LCD = lc + ld + sh, LT = L + sh;
    double ncon_e = 0, ncon_f = 0;
    for (int A = la; A <= la + lb + sh; ++A) ncon_e += fm_ncart(A);
    for (int C = lc; C <= lc + ld + sh; ++C) ncon_f += fm_ncart(C);
    FlopParts fp;
    fp.per_prim = 30.0 + (30.0 + 3.0 * LT) + vrr_flops(LAB, LCD, LT) + 2.0 * ncon_e * ncon_f;
    // HRR targets
    double nf_ket = 0;
    for (int C = lc; C <= lc + ld; ++C) nf_ket += fm_ncart(C);
    double bra_t = 0;
    for (int kb = 1; kb <= lb; ++kb) {
        double na_range = 0;
        for (int A = la; A <= la + lb - kb; ++A) na_range += fm_ncart(A);
        bra_t += na_range * fm_ncart(kb);
    }
    double ket_t = 0;
    for (int kd = 1; kd <= ld; ++kd) {
        double nc_range = 0;
        for (int C = lc; C <= lc + ld - kd; ++C) nc_range += fm_ncart(C);
        ket_t += nc_range * fm_ncart(kd);
    }
    double hrr = 2.0 * (bra_t * nf_ket + ket_t * fm_ncart(la) * fm_ncart(lb));
    // spherical
    int dims[4] = {fm_ncart(la), fm_ncart(lb), fm_ncart(lc), fm_ncart(ld)};
    const int ls[4] = {la, lb, lc, ld};
    double sph = 0;
    if (pure)
        for (int t = 0; t < 4; ++t) {
            if (ls[t] >= 2) {
                double other = 1;
                for (int u = 0; u < 4; ++u)
                    if (u != t) other *= dims[u];
                sph += 2.0 * 8.0 * other;
            }
            if (ls[t] >= 1) dims[t] = 2 * ls[t] + 1;
        }
    double nfq = 1;
    for (int t = 0; t < 4; ++t) nfq *= pure ? (ls[t] >= 1 ? 2 * ls[t] + 1 : 1) : fm_ncart(ls[t]);
    if (gradient)
        fp.per_quartet = 3.0 * (hrr + sph) + 26.0 * nfq;
    else
        fp.per_quartet = hrr + sph + 12.0 * nfq;
    return fp;
}

}  // namespace hfg
