#include "oneint.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace hfscf {

void boys_host(int mmax, double T, double* F) {
    if (T < 1e-15) {
        for (int m = 0; m <= mmax; ++m) F[m] = 1.0 / (2 * m + 1);
        return;
    }
    const double et = std::exp(-T);
    if (T > 35.0 + mmax) {
        F[0] = 0.5 * std::sqrt(M_PI / T) * std::erf(std::sqrt(T));
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - et) / (2 * T);
        return;
    }
    // F_M(T) = e^-T sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)), all terms positive
    long double term = 1.0L / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
        term *= 2.0L * T / (2 * mmax + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-19L) break;
    }
    F[mmax] = (double)(sum * (long double)et);
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2 * T * F[m + 1] + et) / (2 * m + 1);
}

namespace {

struct Comp {
    int l[3];
};

std::vector<Comp> cart_comps(int L) {
    std::vector<Comp> v;
    for (int i = 0; i <= L; ++i)
        for (int j = 0; j <= i; ++j) v.push_back({{L - i, i - j, j}});
    return v;
}

// Hermite expansion coefficients E^{ij}_t of one Cartesian direction.
struct Herm {
    int ni, nj, nt;
    std::vector<double> e;
    double& at(int i, int j, int t) { return e[(i * nj + j) * nt + t]; }
    double get(int i, int j, int t) const {
        if (i < 0 || j < 0 || t < 0 || t > i + j) return 0.0;
        return e[(i * nj + j) * nt + t];
    }
    void build(int imax, int jmax, double p, double xpa, double xpb) {
        ni = imax + 1;
        nj = jmax + 1;
        nt = imax + jmax + 1;
        e.assign((size_t)ni * nj * nt, 0.0);
        const double o2p = 0.5 / p;
        at(0, 0, 0) = 1.0;
        for (int i = 0; i <= imax; ++i)
            for (int j = 0; j <= jmax; ++j) {
                if (i == 0 && j == 0) continue;
                for (int t = 0; t <= i + j; ++t) {
                    double v;
                    if (i > 0)
                        v = o2p * get(i - 1, j, t - 1) + xpa * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1);
                    else
                        v = o2p * get(i, j - 1, t - 1) + xpb * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1);
                    at(i, j, t) = v;
                }
            }
    }
};

// Hermite Coulomb integrals R_tuv(p, PC) for t+u+v <= N.
struct HermR {
    int N;
    std::vector<double> r;  // [t][u][v], N+1 each
    double get(int t, int u, int v) const {
        if (t < 0 || u < 0 || v < 0) return 0.0;
        return r[(t * (N + 1) + u) * (N + 1) + v];
    }
    void build(int n, double p, const double pc[3]) {
        N = n;
        const int D = N + 1;
        std::vector<double> tmp((size_t)D * D * D * D, 0.0);  // [m][t][u][v]
        auto T = [&](int m, int t, int u, int v) -> double& { return tmp[((m * D + t) * D + u) * D + v]; };
        double F[32];
        boys_host(N, p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]), F);
        double f = 1.0;
        for (int m = 0; m <= N; ++m) {
            T(m, 0, 0, 0) = f * F[m];
            f *= -2.0 * p;
        }
        for (int s = 1; s <= N; ++s)          // total t+u+v
            for (int m = 0; m <= N - s; ++m)  // auxiliary order
                for (int t = 0; t <= s; ++t)
                    for (int u = 0; u <= s - t; ++u) {
                        int v = s - t - u;
                        double val;
                        if (t > 0)
                            val = (t > 1 ? (t - 1) * T(m + 1, t - 2, u, v) : 0.0) + pc[0] * T(m + 1, t - 1, u, v);
                        else if (u > 0)
                            val = (u > 1 ? (u - 1) * T(m + 1, t, u - 2, v) : 0.0) + pc[1] * T(m + 1, t, u - 1, v);
                        else
                            val = (v > 1 ? (v - 1) * T(m + 1, t, u, v - 2) : 0.0) + pc[2] * T(m + 1, t, u, v - 1);
                        T(m, t, u, v) = val;
                    }
        r.assign((size_t)D * D * D, 0.0);
        for (int t = 0; t <= N; ++t)
            for (int u = 0; u <= N - t; ++u)
                for (int v = 0; v <= N - t - u; ++v) r[(t * D + u) * D + v] = T(0, t, u, v);
    }
};

struct PrimPair {
    double a, b, p, P[3], PA[3], PB[3], K;  // K = exp(-mu R^2) c_a c_b
};

PrimPair prim_pair(const Shell& sa, int i, const Shell& sb, int j) {
    PrimPair pp;
    pp.a = sa.alpha[i];
    pp.b = sb.alpha[j];
    pp.p = pp.a + pp.b;
    double r2 = 0;
    for (int d = 0; d < 3; ++d) {
        pp.P[d] = (pp.a * sa.r[d] + pp.b * sb.r[d]) / pp.p;
        pp.PA[d] = pp.P[d] - sa.r[d];
        pp.PB[d] = pp.P[d] - sb.r[d];
        r2 += (sa.r[d] - sb.r[d]) * (sa.r[d] - sb.r[d]);
    }
    pp.K = std::exp(-pp.a * pp.b / pp.p * r2) * sa.c[i] * sb.c[j];
    return pp;
}

// out (nsa x nsb) = Ta^T * B (nca x ncb) * Tb  (or copy when Cartesian)
void to_output(const Molecule& mol, const Shell& sa, const Shell& sb, const std::vector<double>& B,
               std::vector<double>& out) {
    const int nca = sa.ncart(), ncb = sb.ncart();
    if (!mol.pure) {
        out = B;
        return;
    }
    const int nsa = sa.nsph(), nsb = sb.nsph();
    auto Ta = ylm_matrix(sa.L), Tb = ylm_matrix(sb.L);
    std::vector<double> tmp((size_t)nsa * ncb, 0.0);
    for (int s = 0; s < nsa; ++s)
        for (int p = 0; p < nca; ++p)
            for (int q = 0; q < ncb; ++q) tmp[s * ncb + q] += Ta[p * nsa + s] * B[p * ncb + q];
    out.assign((size_t)nsa * nsb, 0.0);
    for (int s = 0; s < nsa; ++s)
        for (int t = 0; t < nsb; ++t)
            for (int q = 0; q < ncb; ++q) out[s * nsb + t] += tmp[s * ncb + q] * Tb[q * nsb + t];
}

// Cartesian weight block W_c (nca x ncb) = Ta * W_out * Tb^T from an output-basis matrix W (nbf x nbf).
std::vector<double> weight_cart(const Molecule& mol, int ia, int ib, const double* W) {
    const Shell& sa = mol.shells[ia];
    const Shell& sb = mol.shells[ib];
    const int n = mol.nbf();
    const int oa = mol.out_offset(ia), ob = mol.out_offset(ib);
    const int nca = sa.ncart(), ncb = sb.ncart();
    std::vector<double> Wc((size_t)nca * ncb, 0.0);
    if (!mol.pure) {
        for (int p = 0; p < nca; ++p)
            for (int q = 0; q < ncb; ++q) Wc[p * ncb + q] = W[(oa + p) * n + ob + q];
        return Wc;
    }
    const int nsa = sa.nsph(), nsb = sb.nsph();
    auto Ta = ylm_matrix(sa.L), Tb = ylm_matrix(sb.L);
    std::vector<double> tmp((size_t)nca * nsb, 0.0);
    for (int p = 0; p < nca; ++p)
        for (int s = 0; s < nsa; ++s)
            for (int t = 0; t < nsb; ++t) tmp[p * nsb + t] += Ta[p * nsa + s] * W[(oa + s) * n + ob + t];
    for (int p = 0; p < nca; ++p)
        for (int q = 0; q < ncb; ++q)
            for (int t = 0; t < nsb; ++t) Wc[p * ncb + q] += tmp[p * nsb + t] * Tb[q * nsb + t];
    return Wc;
}

void scatter_block(const Molecule& mol, int ia, int ib, const std::vector<double>& blk, std::vector<double>& M) {
    const int n = mol.nbf();
    const int oa = mol.out_offset(ia), ob = mol.out_offset(ib);
    const int na = mol.out_size(ia), nb = mol.out_size(ib);
    for (int p = 0; p < na; ++p)
        for (int q = 0; q < nb; ++q) {
            M[(oa + p) * n + ob + q] = blk[p * nb + q];
            M[(ob + q) * n + oa + p] = blk[p * nb + q];
        }
}

// 1-D overlap and kinetic pieces from Hermite coefficients.
inline double s1(const Herm& E, double sq, int i, int j) { return (j < 0 || i < 0) ? 0.0 : E.get(i, j, 0) * sq; }
inline double t1(const Herm& E, double sq, double b, int i, int j) {
    return -0.5 * (j * (j - 1) * s1(E, sq, i, j - 2) - 2.0 * b * (2 * j + 1) * s1(E, sq, i, j) +
                   4.0 * b * b * s1(E, sq, i, j + 2));
}

}  // namespace

void overlap_kinetic(const Molecule& mol, std::vector<double>& S, std::vector<double>& T) {
    const int n = mol.nbf();
    const int ns = (int)mol.shells.size();
    S.assign((size_t)n * n, 0.0);
    T.assign((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (int ia = 0; ia < ns; ++ia)
        for (int ib = 0; ib <= ia; ++ib) {
            const Shell& sa = mol.shells[ia];
            const Shell& sb = mol.shells[ib];
            auto ca = cart_comps(sa.L), cb = cart_comps(sb.L);
            std::vector<double> Sb(ca.size() * cb.size(), 0.0), Tb(ca.size() * cb.size(), 0.0);
            Herm E[3];
            for (size_t i = 0; i < sa.alpha.size(); ++i)
                for (size_t j = 0; j < sb.alpha.size(); ++j) {
                    PrimPair pp = prim_pair(sa, (int)i, sb, (int)j);
                    for (int d = 0; d < 3; ++d) E[d].build(sa.L, sb.L + 2, pp.p, pp.PA[d], pp.PB[d]);
                    const double sq = std::sqrt(M_PI / pp.p);
                    for (size_t x = 0; x < ca.size(); ++x)
                        for (size_t y = 0; y < cb.size(); ++y) {
                            const int* a = ca[x].l;
                            const int* b = cb[y].l;
                            double sx = s1(E[0], sq, a[0], b[0]), sy = s1(E[1], sq, a[1], b[1]),
                                   sz = s1(E[2], sq, a[2], b[2]);
                            double tx = t1(E[0], sq, pp.b, a[0], b[0]), ty = t1(E[1], sq, pp.b, a[1], b[1]),
                                   tz = t1(E[2], sq, pp.b, a[2], b[2]);
                            Sb[x * cb.size() + y] += pp.K * sx * sy * sz;
                            Tb[x * cb.size() + y] += pp.K * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
                        }
                }
            std::vector<double> o;
            to_output(mol, sa, sb, Sb, o);
            scatter_block(mol, ia, ib, o, S);
            to_output(mol, sa, sb, Tb, o);
            scatter_block(mol, ia, ib, o, T);
        }
}

void nuclear_attraction(const Molecule& mol, std::vector<double>& V) {
    const int n = mol.nbf();
    const int ns = (int)mol.shells.size();
    V.assign((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (int ia = 0; ia < ns; ++ia)
        for (int ib = 0; ib <= ia; ++ib) {
            const Shell& sa = mol.shells[ia];
            const Shell& sb = mol.shells[ib];
            auto ca = cart_comps(sa.L), cb = cart_comps(sb.L);
            std::vector<double> Vb(ca.size() * cb.size(), 0.0);
            Herm E[3];
            HermR R;
            for (size_t i = 0; i < sa.alpha.size(); ++i)
                for (size_t j = 0; j < sb.alpha.size(); ++j) {
                    PrimPair pp = prim_pair(sa, (int)i, sb, (int)j);
                    for (int d = 0; d < 3; ++d) E[d].build(sa.L, sb.L, pp.p, pp.PA[d], pp.PB[d]);
                    const double pre = 2.0 * M_PI / pp.p * pp.K;
                    for (int c = 0; c < mol.natom(); ++c) {
                        double pc[3] = {pp.P[0] - mol.xyz[c][0], pp.P[1] - mol.xyz[c][1], pp.P[2] - mol.xyz[c][2]};
                        R.build(sa.L + sb.L, pp.p, pc);
                        for (size_t x = 0; x < ca.size(); ++x)
                            for (size_t y = 0; y < cb.size(); ++y) {
                                const int* a = ca[x].l;
                                const int* b = cb[y].l;
                                double acc = 0;
                                for (int t = 0; t <= a[0] + b[0]; ++t)
                                    for (int u = 0; u <= a[1] + b[1]; ++u)
                                        for (int v = 0; v <= a[2] + b[2]; ++v)
                                            acc += E[0].get(a[0], b[0], t) * E[1].get(a[1], b[1], u) *
                                                   E[2].get(a[2], b[2], v) * R.get(t, u, v);
                                Vb[x * cb.size() + y] -= mol.Z[c] * pre * acc;
                            }
                    }
                }
            std::vector<double> o;
            to_output(mol, sa, sb, Vb, o);
            scatter_block(mol, ia, ib, o, V);
        }
}

void oneint_gradient(const Molecule& mol, const double* D, const double* Q, double* grad) {
    const int ns = (int)mol.shells.size();
    const int na = mol.natom();
    std::vector<double> total((size_t)na * 3, 0.0);
#pragma omp parallel
    {
        std::vector<double> g((size_t)na * 3, 0.0);
#pragma omp for schedule(dynamic)
        for (int ia = 0; ia < ns; ++ia)
            for (int ib = 0; ib < ns; ++ib) {
                const Shell& sa = mol.shells[ia];
                const Shell& sb = mol.shells[ib];
                auto ca = cart_comps(sa.L), cb = cart_comps(sb.L);
                const int ncb = (int)cb.size();
                auto WD = weight_cart(mol, ia, ib, D);
                auto WQ = weight_cart(mol, ia, ib, Q);
                Herm E[3];
                HermR R;
                for (size_t i = 0; i < sa.alpha.size(); ++i)
                    for (size_t j = 0; j < sb.alpha.size(); ++j) {
                        PrimPair pp = prim_pair(sa, (int)i, sb, (int)j);
                        for (int d = 0; d < 3; ++d) E[d].build(sa.L + 1, sb.L + 3, pp.p, pp.PA[d], pp.PB[d]);
                        const double sq = std::sqrt(M_PI / pp.p);
                        // overlap and kinetic: d/dA = 2a (a+1) - a_x (a-1); d/dB = -d/dA
                        for (size_t x = 0; x < ca.size(); ++x)
                            for (size_t y = 0; y < cb.size(); ++y) {
                                const double wd = WD[x * ncb + y], wq = WQ[x * ncb + y];
                                if (wd == 0.0 && wq == 0.0) continue;
                                const int* a = ca[x].l;
                                const int* b = cb[y].l;
                                double s[3], t[3], ds[3], dt[3];
                                for (int d = 0; d < 3; ++d) {
                                    s[d] = s1(E[d], sq, a[d], b[d]);
                                    t[d] = t1(E[d], sq, pp.b, a[d], b[d]);
                                    ds[d] = 2.0 * pp.a * s1(E[d], sq, a[d] + 1, b[d]) -
                                            (a[d] > 0 ? a[d] * s1(E[d], sq, a[d] - 1, b[d]) : 0.0);
                                    dt[d] = 2.0 * pp.a * t1(E[d], sq, pp.b, a[d] + 1, b[d]) -
                                            (a[d] > 0 ? a[d] * t1(E[d], sq, pp.b, a[d] - 1, b[d]) : 0.0);
                                }
                                for (int d = 0; d < 3; ++d) {
                                    const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
                                    const double dS = pp.K * ds[d] * s[e1] * s[e2];
                                    const double dT = pp.K * (dt[d] * s[e1] * s[e2] + ds[d] * t[e1] * s[e2] +
                                                              ds[d] * s[e1] * t[e2]);
                                    const double v = -2.0 * wq * dS + 2.0 * wd * dT;
                                    g[sa.atom * 3 + d] += v;
                                    g[sb.atom * 3 + d] -= v;
                                }
                            }
                        // nuclear attraction
                        const double pre = 2.0 * M_PI / pp.p * pp.K;
                        for (int c = 0; c < na; ++c) {
                            double pc[3] = {pp.P[0] - mol.xyz[c][0], pp.P[1] - mol.xyz[c][1], pp.P[2] - mol.xyz[c][2]};
                            R.build(sa.L + sb.L + 1, pp.p, pc);
                            auto vint = [&](const int* a, const int* b) {
                                if (a[0] < 0 || a[1] < 0 || a[2] < 0 || b[0] < 0 || b[1] < 0 || b[2] < 0) return 0.0;
                                double acc = 0;
                                for (int t = 0; t <= a[0] + b[0]; ++t)
                                    for (int u = 0; u <= a[1] + b[1]; ++u)
                                        for (int v = 0; v <= a[2] + b[2]; ++v)
                                            acc += E[0].get(a[0], b[0], t) * E[1].get(a[1], b[1], u) *
                                                   E[2].get(a[2], b[2], v) * R.get(t, u, v);
                                return -mol.Z[c] * pre * acc;
                            };
                            for (size_t x = 0; x < ca.size(); ++x)
                                for (size_t y = 0; y < cb.size(); ++y) {
                                    const double wd = WD[x * ncb + y];
                                    if (wd == 0.0) continue;
                                    for (int d = 0; d < 3; ++d) {
                                        int ap[3] = {ca[x].l[0], ca[x].l[1], ca[x].l[2]};
                                        int bp[3] = {cb[y].l[0], cb[y].l[1], cb[y].l[2]};
                                        int am[3] = {ap[0], ap[1], ap[2]}, bm[3] = {bp[0], bp[1], bp[2]};
                                        const int la = ap[d], lb = bp[d];
                                        ap[d] += 1;
                                        am[d] -= 1;
                                        double dA = 2.0 * pp.a * vint(ap, cb[y].l) - (la ? la * vint(am, cb[y].l) : 0.0);
                                        bp[d] += 1;
                                        bm[d] -= 1;
                                        double dB = 2.0 * pp.b * vint(ca[x].l, bp) - (lb ? lb * vint(ca[x].l, bm) : 0.0);
                                        g[sa.atom * 3 + d] += 2.0 * wd * dA;
                                        g[sb.atom * 3 + d] += 2.0 * wd * dB;
                                        g[c * 3 + d] -= 2.0 * wd * (dA + dB);
                                    }
                                }
                        }
                    }
            }
#pragma omp critical
        for (int k = 0; k < na * 3; ++k) total[k] += g[k];
    }
    for (int k = 0; k < na * 3; ++k) grad[k] += total[k];
}

void nuclear_repulsion_gradient(const Molecule& mol, double* grad) {
    for (int a = 0; a < mol.natom(); ++a)
        for (int b = 0; b < mol.natom(); ++b) {
            if (a == b) continue;
            double d[3], r2 = 0;
            for (int k = 0; k < 3; ++k) {
                d[k] = mol.xyz[a][k] - mol.xyz[b][k];
                r2 += d[k] * d[k];
            }
            const double r3 = r2 * std::sqrt(r2);
            for (int k = 0; k < 3; ++k) grad[a * 3 + k] -= d[k] * mol.Z[a] * mol.Z[b] / r3;
        }
}

}  // namespace hfscf
