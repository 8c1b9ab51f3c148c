// ORACLE — test infrastructure only. This file is a CPU restatement of hfscf's RHF
// two-electron algorithm, used as the parity checker for the GPU library and as the
// timed CPU baseline. Nothing in the product (hfscf_amd/, include/) links or calls it.
//
// What it restates (all citations relative to the hfscf tree):
//   * shell normalisation            src/integrals/hfscf_shell.cpp:44-54
//   * primitive pair data            src/integrals/hfscf_shellpair.cpp:33-60
//   * Boys F_nu via the incomplete gamma function with the T >= 1e-8 floor
//                                    src/math/gamma.cpp:18-135 (series / continued fraction,
//                                    Numerical Recipes 3rd ed. 6.2); ln Gamma from std::lgamma
//                                    (hfscf's gammln cancels against itself in gamm_inc)
//   * chained 1-D Obara-Saika x -> y -> z recursions and per-component seeds/cutoffs
//                                    src/integrals/hfscf_osrecur.hpp:211-312,
//                                    src/integrals/hfscf_oseri.cpp:22-246
//   * Schwarz screen matrix          src/integrals/hfscf_oseri_screen.cpp:22-181
//   * Cartesian -> spherical 4-pass  src/integrals/hfscf_transform.cpp:80-154,
//                                    src/math/solid_harmonics.cpp:16-109
//   * packed ERI index               src/math/hfscf_math.hpp:83-101
//   * stored-ERI loop                src/scfsolver/hfscf_hf_hamiltonian.cpp:67-112
//   * conventional Fock digestion    src/scfsolver/hfscf_hf_solver.cpp:254-286
//   * first-derivative ERIs d/dA = 2a (a+1) - a_x (a-1), D by translation
//                                    src/integrals/hfscf_oserideriv1.cpp:20-374, contracted on
//                                    the fly with the RHF two-particle density of
//                                    src/gradient/hfscf_grad_rhf.cpp:121-141 (no stored tensors)
//   * second-derivative ERIs by applying the shift rule twice (d2/dA_x dA_y on one centre:
//     4a^2 (a+1x+1y) - 2a (a+1x)_y (a+1x-1y) - 2a a_x (a-1x+1y) + a_x (a-1x)_y (a-1x-1y)), D by
//     translation, contracted with the Coulomb and exchange densities of hfscf's
//     calc_scf_hessian_tei (src/hessian/hfscf_hessian_ints.cpp:20-313); hfscf evaluates the
//     same derivatives with its chained recursions (src/integrals/hfscf_oserideriv2qq.cpp:28-749,
//     hfscf_oserideriv2pq.cpp:23-799) and stores them as N^4 tensors per direction pair
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846;
const double kPI25 = 2.0 * std::pow(kPi, 2.5);

// ---------------------------------------------------------------- Boys (gamma.cpp:18-135)
double gser(double a, double x, double gln) {
    double ap = a, del = 1.0 / a, sum = del;
    for (;;) {
        ap += 1.0;
        del *= x / ap;
        sum += del;
        if (std::fabs(del) < std::fabs(sum) * DBL_EPSILON) return sum * std::exp(-x + a * std::log(x) - gln);
    }
}

double gcf(double a, double x, double gln) {
    const double FPMIN = DBL_MIN / DBL_EPSILON;
    double b = x + 1.0 - a, c = 1.0 / FPMIN, d = 1.0 / b, h = d;
    for (int i = 1;; ++i) {
        const double an = -i * (i - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < FPMIN) d = FPMIN;
        c = b + an / c;
        if (std::fabs(c) < FPMIN) c = FPMIN;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) <= DBL_EPSILON) break;
    }
    return std::exp(-x + a * std::log(x) - gln) * h;
}

double boys_ref(double m, double x) {
    x = std::max(std::fabs(x), 1e-8);
    const double a = m + 0.5;
    const double gln = std::lgamma(a);
    if (a >= 100) {
        std::fprintf(stderr, "oracle: Boys order beyond the series/continued-fraction range\n");
        std::abort();
    }
    const double p = (x < a + 1.0) ? gser(a, x, gln) : 1.0 - gcf(a, x, gln);
    return 0.5 * std::pow(x, -m - 0.5) * std::exp(gln) * p;
}

// ---------------------------------------------------------------- tensors
struct T5 {  // (l, k, j, i, n) row-major as hfscf's tensor5d
    int d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0;
    std::vector<double> v;
    void resize(int a, int b, int c, int d, int e) {
        d0 = a; d1 = b; d2 = c; d3 = d; d4 = e;
        v.assign((size_t)a * b * c * d * e, 0.0);
    }
    double& operator()(int l, int k, int j, int i, int n) { return v[((((size_t)l * d1 + k) * d2 + j) * d3 + i) * d4 + n]; }
};

// 1-D VRR on A (via PA), B (via PB) and C (via QC) carrying the auxiliary index n
// (osrecurl123, hfscf_osrecur.hpp:211-254)
void vrr_abc(double PQ, double QC, double PA, double PB, double g2ab, double g2cd, double pr, double qr, int l1, int l2,
             int l3, int nmax, T5& J) {
    if (!nmax) return;
    const double g2 = g2ab + g2cd;
    if (l1)
        for (int n = 0; n < nmax; ++n) J(0, 0, 0, 1, n) = PA * J(0, 0, 0, 0, n) - pr * PQ * J(0, 0, 0, 0, n + 1);
    for (int i = 1; i < l1; ++i)
        for (int n = 0; n < nmax - i; ++n)
            J(0, 0, 0, i + 1, n) = PA * J(0, 0, 0, i, n) - pr * PQ * J(0, 0, 0, i, n + 1) +
                                   (i / g2ab) * (J(0, 0, 0, i - 1, n) - pr * J(0, 0, 0, i - 1, n + 1));
    for (int j = 0; j < l2; ++j)
        for (int i = 0; i <= l1; ++i)
            for (int n = 0; n < nmax - i - j; ++n) {
                double t = PB * J(0, 0, j, i, n) - pr * PQ * J(0, 0, j, i, n + 1);
                if (i) t += (i / g2ab) * (J(0, 0, j, i - 1, n) - pr * J(0, 0, j, i - 1, n + 1));
                if (j) t += (j / g2ab) * (J(0, 0, j - 1, i, n) - pr * J(0, 0, j - 1, i, n + 1));
                J(0, 0, j + 1, i, n) = t;
            }
    for (int k = 0; k < l3; ++k)
        for (int j = 0; j <= l2; ++j)
            for (int i = 0; i <= l1; ++i)
                for (int n = 0; n < nmax - i - j - k; ++n) {
                    double t = QC * J(0, k, j, i, n) + qr * PQ * J(0, k, j, i, n + 1);
                    if (i) t += (i / g2) * J(0, k, j, i - 1, n + 1);
                    if (j) t += (j / g2) * J(0, k, j - 1, i, n + 1);
                    if (k) t += (k / g2cd) * (J(0, k - 1, j, i, n) - qr * J(0, k - 1, j, i, n + 1));
                    J(0, k + 1, j, i, n) = t;
                }
}

// 1-D VRR on A and C only (osrecurl13, hfscf_osrecur.hpp:256-286)
void vrr_ac(double PQ, double QC, double PA, double g2ab, double g2cd, double pr, double qr, int l1, int l3, int nmax,
            T5& J) {
    if (!nmax) return;
    const double g2 = g2ab + g2cd;
    if (l1)
        for (int n = 0; n < nmax; ++n) J(0, 0, 0, 1, n) = PA * J(0, 0, 0, 0, n) - pr * PQ * J(0, 0, 0, 0, n + 1);
    for (int i = 1; i < l1; ++i)
        for (int n = 0; n < nmax - i; ++n)
            J(0, 0, 0, i + 1, n) = PA * J(0, 0, 0, i, n) - pr * PQ * J(0, 0, 0, i, n + 1) +
                                   (i / g2ab) * (J(0, 0, 0, i - 1, n) - pr * J(0, 0, 0, i - 1, n + 1));
    for (int k = 0; k < l3; ++k)
        for (int i = 0; i <= l1; ++i)
            for (int n = 0; n < nmax - i - k; ++n) {
                double t = QC * J(0, k, 0, i, n) + qr * PQ * J(0, k, 0, i, n + 1);
                if (i) t += (i / g2) * J(0, k, 0, i - 1, n + 1);
                if (k) t += (k / g2cd) * (J(0, k - 1, 0, i, n) - qr * J(0, k - 1, 0, i, n + 1));
                J(0, k + 1, 0, i, n) = t;
            }
}

// ---------------------------------------------------------------- shells
struct Shell {
    int L;
    double r[3];
    std::vector<double> a, c;
    int idx, ids;  // Cartesian and spherical offsets
    int atom;
    std::vector<int> lx, ly, lz;
};

struct Pair {
    int s1, s2;
    double AB[3];
    std::vector<double> g, K, P, PA;  // P and PA: 3 per primitive pair
};

double dfac(int n) {
    double r = 1;
    for (int i = n; i > 1; i -= 2) r *= i;
    return r;
}

double binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
double fact(int n) {
    double r = 1;
    for (int i = 2; i <= n; ++i) r *= i;
    return r;
}

// Cartesian -> spherical coefficients, ncart x nsph (solid_harmonics.cpp:16-109)
std::vector<double> ylm(int l) {
    const int nc = (l + 1) * (l + 2) / 2, ns = 2 * l + 1;
    std::vector<double> T((size_t)nc * ns, 0.0);
    const int cosf[4] = {1, 0, -1, 0}, sinf[4] = {0, 1, 0, -1};
    for (int mval = -l; mval <= l; ++mval) {
        const int m = std::abs(mval);
        double pre = std::pow(2.0, -l);
        if (m) pre *= std::sqrt(fact(l - m) * 2.0 / fact(l + m));
        for (int k = 0; k <= (l - m) / 2; ++k) {
            double ff = ((k & 1) ? -1.0 : 1.0) * binom(l, k) * binom(2 * (l - k), l);
            if (m) ff *= fact(l - 2 * k) / fact(l - 2 * k - m);
            ff *= pre;
            for (int a = 0; a <= k; ++a) {
                const double af = binom(k, a) * ff;
                for (int b = 0; b <= a; ++b) {
                    const double f = binom(a, b) * af;
                    const int z = 2 * (b - k) + l - m, y = 2 * (a - b);
                    auto put = [&](int yy, int zz, double w) {
                        const int ii = yy + zz;
                        T[(size_t)(ii * (ii + 1) / 2 + zz) * ns + (l + mval)] += w;
                    };
                    if (mval > 0) {
                        for (int p = 0; p <= m; ++p)
                            if (cosf[(m - p) % 4]) put(y + m - p, z, cosf[(m - p) % 4] * binom(m, p) * f);
                    } else if (m == 0) {
                        put(y, z, f);
                    } else {
                        for (int p = 0; p <= m; ++p)
                            if (sinf[(m - p) % 4]) put(y + m - p, z, sinf[(m - p) % 4] * binom(m, p) * f);
                    }
                }
            }
        }
    }
    return T;
}

struct Oracle {
    bool pure;
    double cutoff, screen_tol;
    bool do_screen;
    int ncart = 0, nsph = 0;
    std::vector<Shell> sh;
    std::vector<Pair> sp;  // n^2 ordered pairs
    std::vector<double> screen;
    std::vector<double> eri;
    std::vector<std::vector<double>> Y;  // per L

    int nbf() const { return pure ? nsph : ncart; }
    int nsh() const { return (int)sh.size(); }
    const Pair& pair(int i, int j) const { return sp[(size_t)i * sh.size() + j]; }
};

void set_pair(Oracle& o, Pair& P, int i, int j) {
    const Shell& A = o.sh[i];
    const Shell& B = o.sh[j];
    P.s1 = i;
    P.s2 = j;
    double r2 = 0;
    for (int d = 0; d < 3; ++d) {
        P.AB[d] = A.r[d] - B.r[d];
        r2 += P.AB[d] * P.AB[d];
    }
    const size_t n = A.a.size() * B.a.size();
    P.g.resize(n);
    P.K.resize(n);
    P.P.resize(3 * n);
    P.PA.resize(3 * n);
    for (size_t x = 0; x < A.a.size(); ++x)
        for (size_t y = 0; y < B.a.size(); ++y) {
            const size_t k = x * B.a.size() + y;
            const double g = A.a[x] + B.a[y];
            P.g[k] = g;
            P.K[k] = std::exp(-A.a[x] * B.a[y] / g * r2) * A.c[x] * B.c[y];
            for (int d = 0; d < 3; ++d) {
                P.P[3 * k + d] = (A.a[x] * A.r[d] + B.a[y] * B.r[d]) / g;
                P.PA[3 * k + d] = P.P[3 * k + d] - A.r[d];
            }
        }
}

inline int64_t index_ij(int64_t i, int64_t j) { return i > j ? (i * (i + 1) >> 1) + j : (j * (j + 1) >> 1) + i; }
inline int64_t index_ijkl(int64_t i, int64_t j, int64_t k, int64_t l) {
    const int64_t ij = index_ij(i, j), kl = index_ij(k, l);
    return ij > kl ? (ij * (ij + 1) >> 1) + kl : (kl * (kl + 1) >> 1) + ij;
}

// Contracted Cartesian quartet by the chained 1-D scheme (hfscf_oseri.cpp:48-200).
// Angular momenta are given explicitly so the same routine also evaluates the shifted
// classes needed for derivatives; wmode selects a per-primitive weight, the product of
// the factors in its 2-bit fields (0: 1, 1: 2 alpha_A, 2: 2 alpha_B, 3: 2 alpha_C). skip_mode: 0 = compute every
// component, 1 = hfscf's Cartesian permutation skips (idx1 > idx2, idx3 > idx4),
// 2 = additionally hfscf's element Schwarz skip and (pq|pq) taken from the screen.
// The per-component seed cutoff uses `seed_cut`; the primitive cutoff `prim_cut`.
struct QuartetWork {
    T5 Jx, Jy, Jz;
};

void cart_quartet(const Oracle& o, const Pair& s12, const Pair& s34, const int L[4], int wmode, int skip_mode,
                  double prim_cut, double seed_cut, bool reset_tmp, double rpq_exact, std::vector<double>& blk,
                  QuartetWork& w) {
    const Shell& A = o.sh[s12.s1];
    const Shell& B = o.sh[s12.s2];
    const Shell& C = o.sh[s34.s1];
    const Shell& D = o.sh[s34.s2];
    const int L1 = L[0], L2 = L[1], L3 = L[2], L4 = L[3];
    auto comps = [](int l, std::vector<int>& x, std::vector<int>& y, std::vector<int>& z) {
        x.clear(); y.clear(); z.clear();
        for (int i = 0; i <= l; ++i)
            for (int j = 0; j <= i; ++j) {
                x.push_back(l - i);
                y.push_back(i - j);
                z.push_back(j);
            }
    };
    std::vector<int> x1, y1, z1, x2, y2, z2, x3, y3, z3, x4, y4, z4;
    comps(L1, x1, y1, z1);
    comps(L2, x2, y2, z2);
    comps(L3, x3, y3, z3);
    comps(L4, x4, y4, z4);
    const int m1 = (int)x1.size(), m2 = (int)x2.size(), m3 = (int)x3.size(), m4 = (int)x4.size();
    blk.assign((size_t)m1 * m2 * m3 * m4, 0.0);
    const int Lt = L1 + L2 + L3 + L4;
    const int nmax = 3 * Lt, uvmax = 2 * Lt, vmax = Lt;
    w.Jx.resize(L4 + 1, L3 + L4 + 1, L2 + 1, L1 + 1, nmax + 1);
    w.Jy.resize(L4 + 1, L3 + L4 + 1, L2 + 1, L1 + 1, uvmax + 1);
    w.Jz.resize(L4 + 1, L3 + L4 + 1, L2 + 1, L1 + L2 + 1, vmax + 1);
    T5& Jx = w.Jx;
    T5& Jy = w.Jy;
    T5& Jz = w.Jz;
    const double* AB = s12.AB;
    const double* CD = s34.AB;
    const size_t nb = B.a.size(), nd = D.a.size();
    for (size_t p1 = 0; p1 < A.a.size(); ++p1)
        for (size_t p2 = 0; p2 < nb; ++p2) {
            const size_t id12 = p1 * nb + p2;
            const double gab = s12.g[id12], gab2 = 2 * gab;
            const double* P = &s12.P[3 * id12];
            const double* PA = &s12.PA[3 * id12];
            double PB[3];
            for (int d = 0; d < 3; ++d) PB[d] = P[d] - B.r[d];
            for (size_t p3 = 0; p3 < C.a.size(); ++p3)
                for (size_t p4 = 0; p4 < nd; ++p4) {
                    const size_t id34 = p3 * nd + p4;
                    const double* Q = &s34.P[3 * id34];
                    const double* QC = &s34.PA[3 * id34];
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    const double rpq2 = PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2];
                    const double gcd = s34.g[id34], gcd2 = 2 * gcd;
                    const double eta = gab * gcd / (gab + gcd);
                    const double pr = eta / gab, qr = eta / gcd;
                    double pfac = s12.K[id12] * s34.K[id34] * kPI25 / (gab * gcd * std::sqrt(gab + gcd));
                    if (std::fabs(pfac) < prim_cut) continue;
                    for (int wm = wmode; wm; wm >>= 2) {  // up to two factors, 2 bits each
                        if ((wm & 3) == 1) pfac *= 2 * A.a[p1];
                        if ((wm & 3) == 2) pfac *= 2 * B.a[p2];
                        if ((wm & 3) == 3) pfac *= 2 * C.a[p3];
                    }
                    const double T = eta * rpq2, eT = std::exp(-T);
                    if (std::fabs(rpq2) > rpq_exact) {
                        Jx(0, 0, 0, 0, nmax) = boys_ref(nmax, T);
                        for (int n = nmax - 1; n >= 0; --n) Jx(0, 0, 0, 0, n) = (1.0 / (2.0 * n + 1.0)) * (2.0 * T * Jx(0, 0, 0, 0, n + 1) + eT);
                    } else {
                        Jx(0, 0, 0, 0, nmax) = 1.0;
                        for (int n = nmax - 1; n >= 0; --n) Jx(0, 0, 0, 0, n) = 1.0 / (2.0 * n + 1.0);
                    }
                    if (nmax == 0) {
                        blk[0] += pfac * Jx(0, 0, 0, 0, 0);
                        continue;
                    }
                    vrr_abc(PQ[0], QC[0], PA[0], PB[0], gab2, gcd2, pr, qr, L1, L2, L3 + L4, nmax, Jx);
                    for (int l = 0; l < L4; ++l)
                        for (int k = 0; k <= L3 + L4; ++k)
                            for (int j = 0; j <= L2; ++j)
                                for (int i = 0; i <= L1; ++i)
                                    for (int n = 0; n <= nmax - i - j - k - l; ++n)
                                        Jx(l + 1, k, j, i, n) = Jx(l, k + 1, j, i, n) + CD[0] * Jx(l, k, j, i, n);
                    for (int ci = 0; ci < m1; ++ci) {
                        const int idx1 = A.idx + ci;
                        for (int cj = 0; cj < m2; ++cj) {
                            const int idx2 = B.idx + cj;
                            if (skip_mode >= 1 && !o.pure && idx1 > idx2) continue;
                            for (int ck = 0; ck < m3; ++ck) {
                                const int idx3 = C.idx + ck;
                                for (int cl = 0; cl < m4; ++cl) {
                                    const int idx4 = D.idx + cl;
                                    if (skip_mode >= 1 && !o.pure && idx3 > idx4) continue;
                                    if (skip_mode == 2) {
                                        if (std::sqrt(std::fabs(o.screen[(size_t)idx1 * o.ncart + idx2] *
                                                                o.screen[(size_t)idx3 * o.ncart + idx4])) < o.screen_tol)
                                            continue;
                                        if (!o.pure && idx1 == idx3 && idx2 == idx4) continue;
                                    }
                                    int nm = y1[ci] + y2[cj] + y3[ck] + y4[cl] + z1[ci] + z2[cj] + z3[ck] + z4[cl];
                                    double tmp = 0;
                                    for (int n = 0; n <= nm; ++n) {
                                        const double v = Jx(x4[cl], x3[ck], x2[cj], x1[ci], n);
                                        Jy(0, 0, 0, 0, n) = v;
                                        tmp += std::fabs(v);
                                    }
                                    if (tmp < seed_cut) continue;
                                    vrr_abc(PQ[1], QC[1], PA[1], PB[1], gab2, gcd2, pr, qr, y1[ci], y2[cj], y3[ck] + y4[cl], nm, Jy);
                                    // HRR on D for the fixed (i, j) of this component (osrecurl4n)
                                    for (int l = 0; l < y4[cl]; ++l)
                                        for (int k = 0; k <= y3[ck] + y4[cl]; ++k)
                                            for (int n = 0; n <= nm - y1[ci] - y2[cj] - l - k; ++n)
                                                Jy(l + 1, k, y2[cj], y1[ci], n) =
                                                    Jy(l, k + 1, y2[cj], y1[ci], n) + CD[1] * Jy(l, k, y2[cj], y1[ci], n);
                                    nm -= y1[ci] + y2[cj] + y3[ck] + y4[cl];
                                    if (reset_tmp) tmp = 0;
                                    for (int n = 0; n <= nm; ++n) {
                                        const double v = Jy(y4[cl], y3[ck], y2[cj], y1[ci], n);
                                        Jz(0, 0, 0, 0, n) = v;
                                        tmp += std::fabs(v);
                                    }
                                    if (tmp < seed_cut) continue;
                                    vrr_ac(PQ[2], QC[2], PA[2], gab2, gcd2, pr, qr, z1[ci] + z2[cj], z3[ck] + z4[cl], nm, Jz);
                                    // HRR on B then D (osrecurl24)
                                    const int i_ = z1[ci], j_ = z2[cj], k_ = z3[ck], l_ = z4[cl];
                                    for (int mm = 0; mm <= k_ + l_; ++mm)
                                        for (int l = 0; l < j_; ++l)
                                            for (int k = 0; k <= i_ + j_; ++k)
                                                Jz(0, mm, l + 1, k, 0) = Jz(0, mm, l, k + 1, 0) + AB[2] * Jz(0, mm, l, k, 0);
                                    for (int l = 0; l < l_; ++l)
                                        for (int k = 0; k <= k_ + l_; ++k)
                                            Jz(l + 1, k, j_, i_, 0) = Jz(l, k + 1, j_, i_, 0) + CD[2] * Jz(l, k, j_, i_, 0);
                                    blk[(((size_t)ci * m2 + cj) * m3 + ck) * m4 + cl] += pfac * Jz(l_, k_, j_, i_, 0);
                                }
                            }
                        }
                    }
                }
        }
}

// 4-pass Cartesian -> spherical transform (hfscf_transform.cpp:80-154)
void to_pure(const Oracle& o, const int L[4], const std::vector<double>& c, std::vector<double>& out) {
    int nc[4], ns[4];
    for (int t = 0; t < 4; ++t) {
        nc[t] = (L[t] + 1) * (L[t] + 2) / 2;
        ns[t] = 2 * L[t] + 1;
    }
    const std::vector<double>* Y[4] = {&o.Y[L[0]], &o.Y[L[1]], &o.Y[L[2]], &o.Y[L[3]]};
    std::vector<double> cur = c;
    int dims[4] = {nc[0], nc[1], nc[2], nc[3]};
    for (int t = 0; t < 4; ++t) {
        int nd[4] = {dims[0], dims[1], dims[2], dims[3]};
        nd[t] = ns[t];
        std::vector<double> nxt((size_t)nd[0] * nd[1] * nd[2] * nd[3], 0.0);
        for (int a = 0; a < dims[0]; ++a)
            for (int b = 0; b < dims[1]; ++b)
                for (int cc = 0; cc < dims[2]; ++cc)
                    for (int d = 0; d < dims[3]; ++d) {
                        const double v = cur[(((size_t)a * dims[1] + b) * dims[2] + cc) * dims[3] + d];
                        if (v == 0.0) continue;
                        int idx[4] = {a, b, cc, d};
                        const int p = idx[t];
                        for (int s = 0; s < ns[t]; ++s) {
                            const double w = (*Y[t])[(size_t)p * ns[t] + s];
                            if (std::fabs(w) < 1e-16) continue;
                            int q[4] = {a, b, cc, d};
                            q[t] = s;
                            nxt[(((size_t)q[0] * nd[1] + q[1]) * nd[2] + q[2]) * nd[3] + q[3]] += w * v;
                        }
                    }
        cur.swap(nxt);
        for (int k = 0; k < 4; ++k) dims[k] = nd[k];
    }
    out.swap(cur);
}

// hfscf: Erios::compute_screen_matrix (hfscf_oseri_screen.cpp:22-181)
void compute_screen(Oracle& o) {
    o.screen.assign((size_t)o.ncart * o.ncart, 0.0);
    const int ns = o.nsh();
#pragma omp parallel
    {
        QuartetWork w;
        std::vector<double> blk;
#pragma omp for schedule(dynamic)
        for (int i = 0; i < ns; ++i)
            for (int j = i; j < ns; ++j) {
                const Pair& P = o.pair(i, j);
                const int L[4] = {o.sh[i].L, o.sh[j].L, o.sh[i].L, o.sh[j].L};
                cart_quartet(o, P, P, L, 0, 1, 1e-16, 1e-12, true, 1e-14, blk, w);
                const int m1 = (L[0] + 1) * (L[0] + 2) / 2, m2 = (L[1] + 1) * (L[1] + 2) / 2;
                for (int p = 0; p < m1; ++p)
                    for (int q = 0; q < m2; ++q) {
                        const int P_ = o.sh[i].idx + p, Q_ = o.sh[j].idx + q;
                        if (P_ <= Q_) {
                            const double v = blk[(((size_t)p * m2 + q) * m1 + p) * m2 + q];
                            o.screen[(size_t)P_ * o.ncart + Q_] = v;
                            o.screen[(size_t)Q_ * o.ncart + P_] = v;
                        }
                    }
            }
    }
}

// hfscf: Erios::compute_contracted_shell_quartet (hfscf_oseri.cpp:22-246), writing eri[]
void quartet_store(Oracle& o, int i, int j, int k, int l, QuartetWork& w) {
    const Pair& s12 = o.pair(i, j);
    const Pair& s34 = o.pair(k, l);
    const Shell &A = o.sh[i], &B = o.sh[j], &C = o.sh[k], &D = o.sh[l];
    const int L[4] = {A.L, B.L, C.L, D.L};
    const int m1 = (A.L + 1) * (A.L + 2) / 2, m2 = (B.L + 1) * (B.L + 2) / 2, m3 = (C.L + 1) * (C.L + 2) / 2,
              m4 = (D.L + 1) * (D.L + 2) / 2;
    if (o.do_screen) {
        double sum = 0;
        for (int p = A.idx; p < A.idx + m1; ++p)
            for (int q = B.idx; q < B.idx + m2; ++q)
                for (int r = C.idx; r < C.idx + m3; ++r)
                    for (int s = D.idx; s < D.idx + m4; ++s)
                        sum += std::sqrt(std::fabs(o.screen[(size_t)p * o.ncart + q] * o.screen[(size_t)r * o.ncart + s]));
        if (sum < o.screen_tol * m1 * m2 * m3 * m4) return;
    }
    std::vector<double> blk;
    cart_quartet(o, s12, s34, L, 0, o.do_screen ? 2 : 1, o.cutoff / 10000.0, o.cutoff, false, 1e-18, blk, w);
    if (!o.pure) {
        for (int p = 0; p < m1; ++p)
            for (int q = 0; q < m2; ++q)
                for (int r = 0; r < m3; ++r)
                    for (int s = 0; s < m4; ++s) {
                        const int P_ = A.idx + p, Q_ = B.idx + q, R_ = C.idx + r, S_ = D.idx + s;
                        if (P_ <= Q_ && R_ <= S_) {
                            double v = blk[(((size_t)p * m2 + q) * m3 + r) * m4 + s];
                            if (o.do_screen && P_ == R_ && Q_ == S_) v = o.screen[(size_t)P_ * o.ncart + Q_];
                            o.eri[index_ijkl(P_, Q_, R_, S_)] = v;
                        }
                    }
    } else {
        std::vector<double> pb;
        if (L[0] + L[1] + L[2] + L[3])
            to_pure(o, L, blk, pb);
        else
            pb = blk;
        const int n1 = 2 * A.L + 1, n2 = 2 * B.L + 1, n3 = 2 * C.L + 1, n4 = 2 * D.L + 1;
        for (int p = 0; p < n1; ++p)
            for (int q = 0; q < n2; ++q)
                for (int r = 0; r < n3; ++r)
                    for (int s = 0; s < n4; ++s) {
                        const int P_ = A.ids + p, Q_ = B.ids + q, R_ = C.ids + r, S_ = D.ids + s;
                        if (P_ <= Q_ && R_ <= S_)
                            o.eri[index_ijkl(P_, Q_, R_, S_)] = pb[(((size_t)p * n2 + q) * n3 + r) * n4 + s];
                    }
    }
}

}  // namespace

extern "C" {

// shells: L, centre (3 per shell, bohr), atom index, nprim, exponents and raw
// (file) contraction coefficients concatenated over shells.
void* orc_create(int nshell, const int* L, const double* xyz, const int* atom, const int* nprim, const double* alpha,
                 const double* coef, int pure, double integral_tol, int do_screen) {
    Oracle* o = new Oracle();
    o->pure = pure != 0;
    o->cutoff = integral_tol;
    o->screen_tol = 100.0 * integral_tol;
    o->do_screen = do_screen != 0;
    int off = 0, idx = 0, ids = 0;
    for (int s = 0; s < nshell; ++s) {
        Shell S;
        S.L = L[s];
        S.atom = atom[s];
        for (int d = 0; d < 3; ++d) S.r[d] = xyz[3 * s + d];
        S.a.assign(alpha + off, alpha + off + nprim[s]);
        std::vector<double> c(coef + off, coef + off + nprim[s]);
        off += nprim[s];
        // normalisation (hfscf_shell.cpp:44-54)
        const double m = S.L + 1.5, m2 = 0.5 * m;
        double sum = 0;
        for (size_t j = 0; j < c.size(); ++j)
            for (size_t k = 0; k < c.size(); ++k)
                sum += c[j] * c[k] * std::pow(S.a[j] * S.a[k], m2) / std::pow(S.a[j] + S.a[k], m);
        const double norm = 1.0 / std::sqrt(sum * std::pow(kPi, 1.5) * dfac(2 * S.L - 1) / std::pow(2.0, S.L));
        S.c.resize(c.size());
        for (size_t j = 0; j < c.size(); ++j) S.c[j] = c[j] * norm * std::pow(S.a[j], m2);
        S.idx = idx;
        S.ids = ids;
        idx += (S.L + 1) * (S.L + 2) / 2;
        ids += 2 * S.L + 1;
        o->sh.push_back(S);
    }
    o->ncart = idx;
    o->nsph = ids;
    o->sp.resize((size_t)nshell * nshell);
    for (int i = 0; i < nshell; ++i)
        for (int j = 0; j < nshell; ++j) set_pair(*o, o->sp[(size_t)i * nshell + j], i, j);
    int lmax = 0;
    for (auto& s : o->sh) lmax = std::max(lmax, s.L);
    for (int l = 0; l <= lmax; ++l) o->Y.push_back(ylm(l));
    if (o->do_screen) compute_screen(*o);
    return o;
}

void orc_destroy(void* h) { delete (Oracle*)h; }

int orc_nbf(void* h) { return ((Oracle*)h)->nbf(); }
int orc_ncart(void* h) { return ((Oracle*)h)->ncart; }

long long orc_num_eri(void* h) {
    const long long n = ((Oracle*)h)->nbf();
    return n * (n + 1) * (n * n + n + 2) / 8;
}

// hfscf: create_repulsion_matrix (hfscf_hf_hamiltonian.cpp:67-112)
void orc_build_eri(void* h) {
    Oracle& o = *(Oracle*)h;
    o.eri.assign((size_t)orc_num_eri(h), 0.0);
    const int ns = o.nsh();
#pragma omp parallel
    {
        QuartetWork w;
#pragma omp for schedule(dynamic)
        for (int i = 0; i < ns; ++i)
            for (int j = i; j < ns; ++j) {
                const int64_t ij = index_ij(i, j);
                for (int k = 0; k < ns; ++k)
                    for (int l = k; l < ns; ++l) {
                        if (ij > index_ij(k, l)) continue;
                        quartet_store(o, i, j, k, l, w);
                    }
            }
    }
}

// Build ERIs only for the shell-quartet rows i in [i0, i1) (bounded CPU timing sample);
// returns the number of shell quartets visited. The (i, j) pairs of the rows are dealt to the
// threads, so that a sample of one row uses every thread as the full build does.
long long orc_build_eri_rows(void* h, int i0, int i1) {
    Oracle& o = *(Oracle*)h;
    if (o.eri.size() != (size_t)orc_num_eri(h)) o.eri.assign((size_t)orc_num_eri(h), 0.0);
    const int ns = o.nsh();
    const long long npair = (long long)std::max(i1 - i0, 0) * ns;
    long long count = 0;
#pragma omp parallel reduction(+ : count)
    {
        QuartetWork w;
#pragma omp for schedule(dynamic)
        for (long long p = 0; p < npair; ++p) {
            const int i = i0 + (int)(p / ns), j = (int)(p % ns);
            if (j < i) continue;
            const int64_t ij = index_ij(i, j);
            for (int k = 0; k < ns; ++k)
                for (int l = k; l < ns; ++l) {
                    if (ij > index_ij(k, l)) continue;
                    quartet_store(o, i, j, k, l, w);
                    ++count;
                }
        }
    }
    return count;
}

void orc_get_eri(void* h, double* out) {
    Oracle& o = *(Oracle*)h;
    std::memcpy(out, o.eri.data(), o.eri.size() * sizeof(double));
}

void orc_get_screen(void* h, double* out) {
    Oracle& o = *(Oracle*)h;
    std::memcpy(out, o.screen.data(), o.screen.size() * sizeof(double));
}

double orc_eri(void* h, int p, int q, int r, int s) { return ((Oracle*)h)->eri[index_ijkl(p, q, r, s)]; }

// hfscf: update_fock_matrix digestion (hfscf_hf_solver.cpp:261-278): F = H + sum D (2(ij|kl) - (ik|jl))
void orc_fock(void* h, const double* H, const double* D, double* F) {
    Oracle& o = *(Oracle*)h;
    const int n = o.nbf();
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            double f = H[(size_t)i * n + j];
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l)
                    f += D[(size_t)k * n + l] * (2.0 * o.eri[index_ijkl(i, j, k, l)] - o.eri[index_ijkl(i, k, j, l)]);
            F[(size_t)i * n + j] = f;
            F[(size_t)j * n + i] = f;
        }
}

void orc_jk(void* h, const double* D, double* J, double* K) {
    Oracle& o = *(Oracle*)h;
    const int n = o.nbf();
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double jj = 0, kk = 0;
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) {
                    jj += D[(size_t)k * n + l] * o.eri[index_ijkl(i, j, k, l)];
                    kk += D[(size_t)k * n + l] * o.eri[index_ijkl(i, k, j, l)];
                }
            J[(size_t)i * n + j] = jj;
            K[(size_t)i * n + j] = kk;
        }
}

// One full shell-quartet block (output basis, all components, no screening), na*nb*nc*nd.
void orc_shell_quartet(void* h, int i, int j, int k, int l, double* out) {
    Oracle& o = *(Oracle*)h;
    QuartetWork w;
    const int L[4] = {o.sh[i].L, o.sh[j].L, o.sh[k].L, o.sh[l].L};
    std::vector<double> blk, pb;
    cart_quartet(o, o.pair(i, j), o.pair(k, l), L, 0, 0, o.cutoff / 10000.0, o.cutoff, false, 1e-18, blk, w);
    if (o.pure && (L[0] + L[1] + L[2] + L[3])) {
        to_pure(o, L, blk, pb);
        std::memcpy(out, pb.data(), pb.size() * sizeof(double));
    } else {
        std::memcpy(out, blk.data(), blk.size() * sizeof(double));
    }
}

// Cartesian derivative blocks d/dA, d/dB, d/dC of one quartet, 9 blocks of
// ncart^4 in the order Ax Ay Az Bx By Bz Cx Cy Cz (hfscf_oserideriv1.cpp:227-352).
void orc_shell_quartet_deriv1(void* h, int i, int j, int k, int l, double* out) {
    Oracle& o = *(Oracle*)h;
    QuartetWork w;
    const int L[4] = {o.sh[i].L, o.sh[j].L, o.sh[k].L, o.sh[l].L};
    int m[4];
    for (int t = 0; t < 4; ++t) m[t] = (L[t] + 1) * (L[t] + 2) / 2;
    const size_t nb = (size_t)m[0] * m[1] * m[2] * m[3];
    std::memset(out, 0, 9 * nb * sizeof(double));
    auto cidx = [](int lx, int ly, int lz) {
        (void)lx;
        const int ii = ly + lz;
        return ii * (ii + 1) / 2 + lz;
    };
    auto comp = [](int l, int c, int e[3]) {
        int ci = 0;
        for (int i = 0; i <= l; ++i)
            for (int j = 0; j <= i; ++j, ++ci)
                if (ci == c) {
                    e[0] = l - i;
                    e[1] = i - j;
                    e[2] = j;
                    return;
                }
    };
    for (int centre = 0; centre < 3; ++centre) {
        int Lp[4] = {L[0], L[1], L[2], L[3]};
        Lp[centre] += 1;
        std::vector<double> up, dn;
        cart_quartet(o, o.pair(i, j), o.pair(k, l), Lp, centre + 1, 0, o.cutoff / 10000.0, 0.0, false, 1e-18, up, w);
        int Lm[4] = {L[0], L[1], L[2], L[3]};
        const bool has_dn = L[centre] > 0;
        if (has_dn) {
            Lm[centre] -= 1;
            cart_quartet(o, o.pair(i, j), o.pair(k, l), Lm, 0, 0, o.cutoff / 10000.0, 0.0, false, 1e-18, dn, w);
        }
        int mp[4], mm[4];
        for (int t = 0; t < 4; ++t) {
            mp[t] = (Lp[t] + 1) * (Lp[t] + 2) / 2;
            mm[t] = (Lm[t] + 1) * (Lm[t] + 2) / 2;
        }
        for (int a = 0; a < m[0]; ++a)
            for (int b = 0; b < m[1]; ++b)
                for (int c = 0; c < m[2]; ++c)
                    for (int d = 0; d < m[3]; ++d) {
                        int ci[4] = {a, b, c, d};
                        int e[3];
                        comp(L[centre], ci[centre], e);
                        for (int dir = 0; dir < 3; ++dir) {
                            int ep[3] = {e[0], e[1], e[2]};
                            ep[dir] += 1;
                            int cp[4] = {a, b, c, d};
                            cp[centre] = cidx(ep[0], ep[1], ep[2]);
                            double v = up[(((size_t)cp[0] * mp[1] + cp[1]) * mp[2] + cp[2]) * mp[3] + cp[3]];
                            if (has_dn && e[dir] > 0) {
                                int em[3] = {e[0], e[1], e[2]};
                                em[dir] -= 1;
                                int cm[4] = {a, b, c, d};
                                cm[centre] = cidx(em[0], em[1], em[2]);
                                v -= e[dir] * dn[(((size_t)cm[0] * mm[1] + cm[1]) * mm[2] + cm[2]) * mm[3] + cm[3]];
                            }
                            out[(size_t)(centre * 3 + dir) * nb + (((size_t)a * m[1] + b) * m[2] + c) * m[3] + d] = v;
                        }
                    }
    }
}

void grad2e_w(void* h, const double* D, int natom, double* grad, double wc, double wx);

// Two-electron RHF gradient sum_pqrs [2 D_pq D_rs - D_pr D_qs] (pq|rs)', contracted on the
// fly over canonical shell quartets (natom x 3 added into grad).
void orc_grad2e(void* h, const double* D, int natom, double* grad) { grad2e_w(h, D, natom, grad, 1.0, 1.0); }

// The same with the Coulomb part weighted by wc and the exchange part by wx (wc = 1, wx = 0:
// hfscf's gradCoulomb; wc = 0, wx = 1: its gradExchange, hfscf_grad_rhf.cpp:140-141).
void orc_grad2e_parts(void* h, const double* D, int natom, double* gradC, double* gradX) {
    grad2e_w(h, D, natom, gradC, 1.0, 0.0);
    grad2e_w(h, D, natom, gradX, 0.0, 1.0);
}

}  // extern "C"

void grad2e_w(void* h, const double* D, int natom, double* grad, double wc, double wx) {
    Oracle& o = *(Oracle*)h;
    const int ns = o.nsh();
    const int n = o.nbf();
    std::vector<double> total((size_t)natom * 3, 0.0);
#pragma omp parallel
    {
        std::vector<double> g((size_t)natom * 3, 0.0);
        std::vector<double> der, pblk, cblk;
#pragma omp for schedule(dynamic)
        for (int a = 0; a < ns; ++a)
            for (int b = 0; b <= a; ++b)
                for (int c = 0; c <= a; ++c)
                    for (int d = 0; d <= c; ++d) {
                        if (index_ij(a, b) < index_ij(c, d)) continue;
                        double deg = 8.0;
                        if (a == b) deg *= 0.5;
                        if (c == d) deg *= 0.5;
                        if (a == c && b == d) deg *= 0.5;
                        const int L[4] = {o.sh[a].L, o.sh[b].L, o.sh[c].L, o.sh[d].L};
                        int m[4], no[4], off[4];
                        for (int t = 0; t < 4; ++t) {
                            m[t] = (L[t] + 1) * (L[t] + 2) / 2;
                            no[t] = o.pure ? 2 * L[t] + 1 : m[t];
                        }
                        off[0] = o.pure ? o.sh[a].ids : o.sh[a].idx;
                        off[1] = o.pure ? o.sh[b].ids : o.sh[b].idx;
                        off[2] = o.pure ? o.sh[c].ids : o.sh[c].idx;
                        off[3] = o.pure ? o.sh[d].ids : o.sh[d].idx;
                        const size_t nb = (size_t)m[0] * m[1] * m[2] * m[3];
                        const size_t nbo = (size_t)no[0] * no[1] * no[2] * no[3];
                        der.resize(9 * nb);
                        orc_shell_quartet_deriv1(h, a, b, c, d, der.data());
                        const int atoms[4] = {o.sh[a].atom, o.sh[b].atom, o.sh[c].atom, o.sh[d].atom};
                        for (int comp = 0; comp < 9; ++comp) {
                            const double* blk = der.data() + comp * nb;
                            const double* use = blk;
                            if (o.pure && (L[0] + L[1] + L[2] + L[3])) {
                                cblk.assign(blk, blk + nb);
                                to_pure(o, L, cblk, pblk);
                                use = pblk.data();
                            }
                            double s = 0;
                            for (int p = 0; p < no[0]; ++p)
                                for (int q = 0; q < no[1]; ++q)
                                    for (int r = 0; r < no[2]; ++r)
                                        for (int t = 0; t < no[3]; ++t) {
                                            const int P_ = off[0] + p, Q_ = off[1] + q, R_ = off[2] + r, S_ = off[3] + t;
                                            const double gam = wc * 2.0 * D[(size_t)P_ * n + Q_] * D[(size_t)R_ * n + S_] -
                                                               wx * 0.5 * (D[(size_t)P_ * n + R_] * D[(size_t)Q_ * n + S_] +
                                                                           D[(size_t)P_ * n + S_] * D[(size_t)Q_ * n + R_]);
                                            s += gam * use[(((size_t)p * no[1] + q) * no[2] + r) * no[3] + t];
                                        }
                            (void)nbo;
                            const int centre = comp / 3, dir = comp % 3;
                            g[atoms[centre] * 3 + dir] += deg * s;
                            g[atoms[3] * 3 + dir] -= deg * s;  // d/dD = -(d/dA + d/dB + d/dC)
                        }
                    }
#pragma omp critical
        for (size_t k = 0; k < total.size(); ++k) total[k] += g[k];
    }
    for (size_t k = 0; k < total.size(); ++k) grad[k] += total[k];
}

// Cartesian second-derivative blocks of one quartet: 45 blocks of ncart^4, one per coordinate
// pair (X1 <= X2) of X = 3 * centre + direction over the centres A, B, C, in the order
// (0,0) (0,1) .. (0,8) (1,1) .. (8,8).
void quartet_deriv2(Oracle& o, int i, int j, int k, int l, std::vector<double>& out) {
    QuartetWork w;
    const int L[4] = {o.sh[i].L, o.sh[j].L, o.sh[k].L, o.sh[l].L};
    int m[4];
    for (int t = 0; t < 4; ++t) m[t] = (L[t] + 1) * (L[t] + 2) / 2;
    const size_t nb = (size_t)m[0] * m[1] * m[2] * m[3];
    out.assign(45 * nb, 0.0);
    auto cidx = [](const int e[3]) {
        const int ii = e[1] + e[2];
        return ii * (ii + 1) / 2 + e[2];
    };
    auto comp = [](int l, int c, int e[3]) {
        int ci = 0;
        for (int a = 0; a <= l; ++a)
            for (int b = 0; b <= a; ++b, ++ci)
                if (ci == c) {
                    e[0] = l - a;
                    e[1] = a - b;
                    e[2] = b;
                    return;
                }
    };
    // shifted classes by (angular momentum shifts, weight mode), evaluated once per quartet
    std::map<int, std::vector<double>> cache;
    auto cls = [&](const int dL[4], int wmode) -> const std::vector<double>& {
        int key = wmode;
        for (int t = 0; t < 4; ++t) key = key * 5 + (dL[t] + 2);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
        int Lp[4];
        for (int t = 0; t < 4; ++t) Lp[t] = L[t] + dL[t];
        std::vector<double>& v = cache[key];
        cart_quartet(o, o.pair(i, j), o.pair(k, l), Lp, wmode, 0, o.cutoff / 10000.0, 0.0, false, 1e-18, v, w);
        return v;
    };
    auto at = [&](const std::vector<double>& v, const int dL[4], const int c4[4]) {
        int mp[4];
        for (int t = 0; t < 4; ++t) mp[t] = (L[t] + dL[t] + 1) * (L[t] + dL[t] + 2) / 2;
        return v[(((size_t)c4[0] * mp[1] + c4[1]) * mp[2] + c4[2]) * mp[3] + c4[3]];
    };
    int pidx = 0;
    for (int X1 = 0; X1 < 9; ++X1)
        for (int X2 = X1; X2 < 9; ++X2, ++pidx) {
            const int c1 = X1 / 3, d1 = X1 % 3, c2 = X2 / 3, d2 = X2 % 3;
            for (int a = 0; a < m[0]; ++a)
                for (int b = 0; b < m[1]; ++b)
                    for (int c = 0; c < m[2]; ++c)
                        for (int d = 0; d < m[3]; ++d) {
                            const int ci[4] = {a, b, c, d};
                            double val = 0.0;
                            if (c1 != c2) {
                                // one shift on each of two centres: (2a+ - e-) x (2b+ - e-)
                                int e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
                                comp(L[c1], ci[c1], e1);
                                comp(L[c2], ci[c2], e2);
                                for (int s1 = 1; s1 >= -1; s1 -= 2)
                                    for (int s2 = 1; s2 >= -1; s2 -= 2) {
                                        const double f = (s1 > 0 ? 1.0 : -e1[d1]) * (s2 > 0 ? 1.0 : -e2[d2]);
                                        if (f == 0.0) continue;
                                        int dL[4] = {0, 0, 0, 0};
                                        dL[c1] = s1;
                                        dL[c2] = s2;
                                        const int wmode = (s1 > 0 ? c1 + 1 : 0) | ((s2 > 0 ? c2 + 1 : 0) << 2);
                                        int f1[3] = {e1[0], e1[1], e1[2]}, f2[3] = {e2[0], e2[1], e2[2]};
                                        f1[d1] += s1;
                                        f2[d2] += s2;
                                        int c4[4] = {a, b, c, d};
                                        c4[c1] = cidx(f1);
                                        c4[c2] = cidx(f2);
                                        val += f * at(cls(dL, wmode), dL, c4);
                                    }
                            } else {
                                // two shifts on one centre
                                int e[3] = {0, 0, 0};
                                comp(L[c1], ci[c1], e);
                                for (int s1 = 1; s1 >= -1; s1 -= 2)
                                    for (int s2 = 1; s2 >= -1; s2 -= 2) {
                                        int f1[3] = {e[0], e[1], e[2]};
                                        double f;
                                        if (s1 > 0) {
                                            f1[d1] += 1;
                                            f = s2 > 0 ? 1.0 : -(double)f1[d2];
                                        } else {
                                            f = -(double)e[d1];
                                            f1[d1] -= 1;
                                            if (s2 < 0) f *= -(double)f1[d2];
                                        }
                                        if (f == 0.0) continue;
                                        f1[d2] += s2;
                                        int dL[4] = {0, 0, 0, 0};
                                        dL[c1] = s1 + s2;
                                        const int wmode = (s1 > 0 ? c1 + 1 : 0) | ((s2 > 0 ? c1 + 1 : 0) << 2);
                                        int c4[4] = {a, b, c, d};
                                        c4[c1] = cidx(f1);
                                        val += f * at(cls(dL, wmode), dL, c4);
                                    }
                            }
                            out[(size_t)pidx * nb + (((size_t)a * m[1] + b) * m[2] + c) * m[3] + d] = val;
                        }
        }
}

extern "C" {

void orc_shell_quartet_deriv2(void* h, int i, int j, int k, int l, double* out) {
    std::vector<double> v;
    quartet_deriv2(*(Oracle*)h, i, j, k, l, v);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
}

// Two-electron part of the RHF Hessian with the density held fixed, split as hfscf's
// m_hessianC = sum D_ij D_kl 2 (ij|kl)'' and m_hessianEx = -sum D_ij D_kl (ik|jl)''
// (hfscf_hessian_ints.cpp:20-313), both (3 natom) x (3 natom), added into HC and HX. Canonical
// shell quartets with their degeneracy, the 8-fold symmetrised densities of the gradient, the
// D-centre rows and columns from translational invariance (d/dD = -(d/dA + d/dB + d/dC)).
void orc_hess2e(void* h, const double* D, int natom, double* HC, double* HX) {
    Oracle& o = *(Oracle*)h;
    const int ns = o.nsh();
    const int n = o.nbf();
    const int n3 = 3 * natom;
    std::vector<double> totC((size_t)n3 * n3, 0.0), totX((size_t)n3 * n3, 0.0);
#pragma omp parallel
    {
        std::vector<double> hc((size_t)n3 * n3, 0.0), hx((size_t)n3 * n3, 0.0);
        std::vector<double> der, pblk, cblk;
#pragma omp for schedule(dynamic)
        for (int a = 0; a < ns; ++a)
            for (int b = 0; b <= a; ++b)
                for (int c = 0; c <= a; ++c)
                    for (int d = 0; d <= c; ++d) {
                        if (index_ij(a, b) < index_ij(c, d)) continue;
                        double deg = 8.0;
                        if (a == b) deg *= 0.5;
                        if (c == d) deg *= 0.5;
                        if (a == c && b == d) deg *= 0.5;
                        const int L[4] = {o.sh[a].L, o.sh[b].L, o.sh[c].L, o.sh[d].L};
                        int m[4], no[4], off[4];
                        for (int t = 0; t < 4; ++t) {
                            m[t] = (L[t] + 1) * (L[t] + 2) / 2;
                            no[t] = o.pure ? 2 * L[t] + 1 : m[t];
                        }
                        const int sh4[4] = {a, b, c, d};
                        for (int t = 0; t < 4; ++t) off[t] = o.pure ? o.sh[sh4[t]].ids : o.sh[sh4[t]].idx;
                        const size_t nb = (size_t)m[0] * m[1] * m[2] * m[3];
                        quartet_deriv2(o, a, b, c, d, der);
                        double MC[9][9], MX[9][9];
                        int pidx = 0;
                        for (int X1 = 0; X1 < 9; ++X1)
                            for (int X2 = X1; X2 < 9; ++X2, ++pidx) {
                                const double* blk = der.data() + pidx * nb;
                                const double* use = blk;
                                if (o.pure && (L[0] + L[1] + L[2] + L[3])) {
                                    cblk.assign(blk, blk + nb);
                                    to_pure(o, L, cblk, pblk);
                                    use = pblk.data();
                                }
                                double sc = 0, sx = 0;
                                for (int p = 0; p < no[0]; ++p)
                                    for (int q = 0; q < no[1]; ++q)
                                        for (int r = 0; r < no[2]; ++r)
                                            for (int t = 0; t < no[3]; ++t) {
                                                const int P_ = off[0] + p, Q_ = off[1] + q, R_ = off[2] + r, S_ = off[3] + t;
                                                const double v = use[(((size_t)p * no[1] + q) * no[2] + r) * no[3] + t];
                                                sc += 2.0 * D[(size_t)P_ * n + Q_] * D[(size_t)R_ * n + S_] * v;
                                                sx -= 0.5 * (D[(size_t)P_ * n + R_] * D[(size_t)Q_ * n + S_] +
                                                             D[(size_t)P_ * n + S_] * D[(size_t)Q_ * n + R_]) * v;
                                            }
                                MC[X1][X2] = MC[X2][X1] = deg * sc;
                                MX[X1][X2] = MX[X2][X1] = deg * sx;
                            }
                        // 12 x 12 over A, B, C, D: T M T^T with the D rows -(A + B + C)
                        const int atoms[4] = {o.sh[a].atom, o.sh[b].atom, o.sh[c].atom, o.sh[d].atom};
                        auto tcoef = [](int Y, int X) {  // T[Y][X], Y over 12, X over 9
                            if (Y < 9) return Y == X ? 1.0 : 0.0;
                            return (X % 3 == Y % 3) ? -1.0 : 0.0;
                        };
                        for (int Y1 = 0; Y1 < 12; ++Y1)
                            for (int Y2 = 0; Y2 < 12; ++Y2) {
                                double vc = 0, vx = 0;
                                for (int X1 = 0; X1 < 9; ++X1) {
                                    const double t1 = tcoef(Y1, X1);
                                    if (t1 == 0.0) continue;
                                    for (int X2 = 0; X2 < 9; ++X2) {
                                        const double t2 = tcoef(Y2, X2);
                                        if (t2 == 0.0) continue;
                                        vc += t1 * t2 * MC[X1][X2];
                                        vx += t1 * t2 * MX[X1][X2];
                                    }
                                }
                                const size_t r = (size_t)atoms[Y1 / 3] * 3 + Y1 % 3, col = (size_t)atoms[Y2 / 3] * 3 + Y2 % 3;
                                hc[r * n3 + col] += vc;
                                hx[r * n3 + col] += vx;
                            }
                    }
#pragma omp critical
        for (size_t q = 0; q < totC.size(); ++q) {
            totC[q] += hc[q];
            totX[q] += hx[q];
        }
    }
    for (size_t q = 0; q < totC.size(); ++q) {
        HC[q] += totC[q];
        HX[q] += totX[q];
    }
}

double orc_boys(int m, double T) { return boys_ref(m, T); }

// OpenMP threads of the oracle's parallel loops (timed CPU baseline); returns the previous count
int orc_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

}  // extern "C"
