// One-electron integrals on the GPU: overlap S, kinetic T and nuclear attraction V of a single
// molecule, and their first-derivative contractions for the RHF gradient. They replace what hfscf
// computes on the CPU in OSOverlap, OSKinetic and OSNuclear and their first derivatives (hfscf:
// src/integrals/hfscf_osoverlap.cpp:10, hfscf_osnuclear.cpp:17-83, hfscf_osoverlapderiv1.cpp:8,
// hfscf_oskineticderiv1.cpp:8, hfscf_osnuclearderiv1.cpp:15), with McMurchie-Davidson Hermite
// expansions: S and T from the 1-D Hermite coefficients E^{ij}_0, V from E^{ij}_t and the Hermite
// Coulomb integrals R_tuv (Boys values from the device table), derivatives by the shifts
// d/dA_x |a> = 2 alpha |a + 1_x> - a_x |a - 1_x>.
//
// One lane per shell pair for S and T, per shell pair and nucleus for V (added into V with
// atomics), and per ordered shell pair and part (the overlap and kinetic terms, or one nucleus)
// for the gradient terms; angular momenta up to 2, so the derivative shifts reach 3 on the bra and,
// for T, 5 on the ket. This is O(N^2 natom) work next to the O(N^4) two-electron build, so the
// kernels keep their tables in per-lane arrays sized for the largest shells rather than
// specialising per class.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "boys.hpp"
#include "context.hpp"
#include "coop.hpp"
#include "oneint.hpp"

namespace hfg {

namespace {

constexpr int kOneMaxL = 2;                   // largest shell angular momentum
constexpr int kRN = 2 * kOneMaxL + 1;         // largest total order of R_tuv (gradient)
constexpr int kRSize = (kRN + 1) * (kRN + 2) * (kRN + 3) / 6;  // R_tuv with t + u + v <= kRN

// Cartesian exponents of component k of a shell of angular momentum L (hfscf's order: xx, xy, xz,
// yy, yz, zz; hfscf: src/integrals/hfscf_shell.cpp:27-40).
__device__ inline void cart_lmn(int L, int k, int* l) {
    int n = 0;
    for (int i = 0; i <= L; ++i)
        for (int j = 0; j <= i; ++j, ++n)
            if (n == k) {
                l[0] = L - i;
                l[1] = i - j;
                l[2] = j;
                return;
            }
}

__device__ inline int ncart1(int L) { return (L + 1) * (L + 2) / 2; }

// Hermite expansion coefficients E^{ij}_t of one direction, i < ni <= EI, j < nj <= EJ: the
// integrals need EI = Lmax + 1, EJ = Lmax + 3 (kinetic: ket + 2), the gradient one more of each.
template <int EI, int EJ>
struct Herm1 {
    static constexpr int kEj = EJ, kEt = EI + EJ - 1;  // t <= i + j
    double e[EI * EJ * kEt];
    int ni, nj;
    __device__ double get(int i, int j, int t) const {
        if (i < 0 || j < 0 || t < 0 || t > i + j || i >= ni || j >= nj) return 0.0;
        return e[(i * kEj + j) * kEt + t];
    }
    __device__ void build(int imax, int jmax, double p, double xpa, double xpb) {
        ni = imax + 1;
        nj = jmax + 1;
        const double o2p = 0.5 / p;
        for (int i = 0; i <= imax; ++i)
            for (int j = 0; j <= jmax; ++j)
                for (int t = 0; t <= i + j; ++t) {
                    double v;
                    if (i == 0 && j == 0)
                        v = 1.0;
                    else if (i > 0)
                        v = o2p * get_partial(i - 1, j, t - 1) + xpa * get_partial(i - 1, j, t) +
                            (t + 1) * get_partial(i - 1, j, t + 1);
                    else
                        v = o2p * get_partial(i, j - 1, t - 1) + xpb * get_partial(i, j - 1, t) +
                            (t + 1) * get_partial(i, j - 1, t + 1);
                    e[(i * kEj + j) * kEt + t] = v;
                }
    }
    // entries already built (the recursion reads only lower i, or lower j at i = 0)
    __device__ double get_partial(int i, int j, int t) const {
        if (t < 0 || t > i + j) return 0.0;
        return e[(i * kEj + j) * kEt + t];
    }
};

__device__ inline int ridx(int t, int u, int v) {
    const int s = t + u + v;
    return s * (s + 1) * (s + 2) / 6 + (s - t) * (s - t + 1) / 2 + v;
}

// Boys values F_0..F_N(T) scaled by `scale`, N <= kRN, from the device table.
__device__ inline void boys_n(int N, double T, const double* tab, double* F) {
    switch (N) {
        case 0: boys_eval<0, double>(T, tab, 1.0, F); break;
        case 1: boys_eval<1, double>(T, tab, 1.0, F); break;
        case 2: boys_eval<2, double>(T, tab, 1.0, F); break;
        case 3: boys_eval<3, double>(T, tab, 1.0, F); break;
        case 4: boys_eval<4, double>(T, tab, 1.0, F); break;
        default: boys_eval<kRN, double>(T, tab, 1.0, F); break;
    }
}

// Hermite Coulomb integrals R_tuv(p, PC), t + u + v <= N, by the recursion over the auxiliary
// order from R^(m)_000 = (-2p)^m F_m(p |PC|^2), one order at a time.
__device__ inline void hermite_r(int N, double p, const double* pc, const double* tab, double* R) {
    double F[kRN + 1];
    boys_n(N, p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]), tab, F);
    double prev[kRSize], cur[kRSize];
    double f = 1.0;
    for (int m = 0; m < N; ++m) f *= -2.0 * p;
    prev[0] = f * F[N];  // order N: only R_000
    for (int m = N - 1; m >= 0; --m) {
        f /= -2.0 * p;
        cur[0] = f * F[m];
        const int smax = N - m;
        for (int s = 1; s <= smax; ++s)
            for (int t = 0; t <= s; ++t)
                for (int u = 0; u <= s - t; ++u) {
                    const int v = s - t - u;
                    double val;
                    if (t > 0)
                        val = (t > 1 ? (t - 1) * prev[ridx(t - 2, u, v)] : 0.0) + pc[0] * prev[ridx(t - 1, u, v)];
                    else if (u > 0)
                        val = (u > 1 ? (u - 1) * prev[ridx(t, u - 2, v)] : 0.0) + pc[1] * prev[ridx(t, u - 1, v)];
                    else
                        val = (v > 1 ? (v - 1) * prev[ridx(t, u, v - 2)] : 0.0) + pc[2] * prev[ridx(t, u, v - 1)];
                    cur[ridx(t, u, v)] = val;
                }
        for (int k = 0; k < (smax + 1) * (smax + 2) * (smax + 3) / 6; ++k) prev[k] = cur[k];
    }
    for (int k = 0; k < (N + 1) * (N + 2) * (N + 3) / 6; ++k) R[k] = prev[k];
}

struct PrimPairDev {
    double a, b, p, P[3], PA[3], PB[3], K;  // K = exp(-ab/p |AB|^2) c_a c_b
};

__device__ inline PrimPairDev prim_pair_dev(const OneArgs& g, int sa, int i, int sb, int j) {
    const OneShell A = g.sh[sa], B = g.sh[sb];
    const double* ra = g.xyz + 3 * A.atom;
    const double* rb = g.xyz + 3 * B.atom;
    PrimPairDev pp;
    pp.a = g.alpha[A.po + i];
    pp.b = g.alpha[B.po + j];
    pp.p = pp.a + pp.b;
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        pp.P[d] = (pp.a * ra[d] + pp.b * rb[d]) / pp.p;
        pp.PA[d] = pp.P[d] - ra[d];
        pp.PB[d] = pp.P[d] - rb[d];
        r2 += (ra[d] - rb[d]) * (ra[d] - rb[d]);
    }
    pp.K = exp(-pp.a * pp.b / pp.p * r2) * g.coef[A.po + i] * g.coef[B.po + j];
    return pp;
}

template <class E1>
__device__ inline double s1(const E1& E, double sq, int i, int j) { return (i < 0 || j < 0) ? 0.0 : E.get(i, j, 0) * sq; }
template <class E1>
__device__ inline double t1(const E1& E, double sq, double b, int i, int j) {
    return -0.5 * (j * (j - 1) * s1(E, sq, i, j - 2) - 2.0 * b * (2 * j + 1) * s1(E, sq, i, j) +
                   4.0 * b * b * s1(E, sq, i, j + 2));
}

// Output-basis block out (na x nb) of a Cartesian block B (nca x ncb): out = Ta^T B Tb (spherical,
// hfscf's order m = -l..l) or B itself.
__device__ inline void block_out(int La, int Lb, bool pure, const double* B, double* out) {
    const int nca = ncart1(La), ncb = ncart1(Lb);
    if (!pure) {
        for (int k = 0; k < nca * ncb; ++k) out[k] = B[k];
        return;
    }
    const int nsa = 2 * La + 1, nsb = 2 * Lb + 1;
    for (int s = 0; s < nsa; ++s) {
        const SphTerms ta = sph_terms(La, s);
        for (int t = 0; t < nsb; ++t) {
            const SphTerms tb = sph_terms(Lb, t);
            double v = 0.0;
            for (int x = 0; x < ta.n; ++x)
                for (int y = 0; y < tb.n; ++y) v += ta.w[x] * tb.w[y] * B[ta.c[x] * ncb + tb.c[y]];
            out[s * nsb + t] = v;
        }
    }
}

// Cartesian weight block Wc (nca x ncb) = Ta W_ab Tb^T of the output-basis matrix W (n x n).
__device__ inline void weight_cart(int La, int Lb, bool pure, const double* W, int n, int oa, int ob, double* Wc) {
    const int nca = ncart1(La), ncb = ncart1(Lb);
    if (!pure) {
        for (int p = 0; p < nca; ++p)
            for (int q = 0; q < ncb; ++q) Wc[p * ncb + q] = W[(size_t)(oa + p) * n + ob + q];
        return;
    }
    for (int k = 0; k < nca * ncb; ++k) Wc[k] = 0.0;
    for (int s = 0; s < 2 * La + 1; ++s) {
        const SphTerms ta = sph_terms(La, s);
        for (int t = 0; t < 2 * Lb + 1; ++t) {
            const SphTerms tb = sph_terms(Lb, t);
            const double w = W[(size_t)(oa + s) * n + ob + t];
            for (int x = 0; x < ta.n; ++x)
                for (int y = 0; y < tb.n; ++y) Wc[ta.c[x] * ncb + tb.c[y]] += ta.w[x] * tb.w[y] * w;
        }
    }
}

__device__ inline void scatter(const OneArgs& g, int sa, int sb, const double* blk, double* M) {
    const OneShell A = g.sh[sa], B = g.sh[sb];
    const int na = g.pure ? 2 * A.L + 1 : ncart1(A.L), nb = g.pure ? 2 * B.L + 1 : ncart1(B.L);
    for (int p = 0; p < na; ++p)
        for (int q = 0; q < nb; ++q) {
            M[(size_t)(A.ao + p) * g.n + B.ao + q] = blk[p * nb + q];
            M[(size_t)(B.ao + q) * g.n + A.ao + p] = blk[p * nb + q];
        }
}

constexpr int kBlk = 36;  // 6 x 6 Cartesian components of two d shells

// Adds block blk (na x nb) of shell pair (sa, sb) into M with atomics: at (a, b) and, for two
// different shells, at (b, a) (the block of a shell with itself already covers both).
__device__ inline void scatter_add(const OneArgs& g, int sa, int sb, const double* blk, double* M) {
    const OneShell A = g.sh[sa], B = g.sh[sb];
    const int na = g.pure ? 2 * A.L + 1 : ncart1(A.L), nb = g.pure ? 2 * B.L + 1 : ncart1(B.L);
    for (int p = 0; p < na; ++p)
        for (int q = 0; q < nb; ++q) {
            atomicAdd(&M[(size_t)(A.ao + p) * g.n + B.ao + q], blk[p * nb + q]);
            if (sa != sb) atomicAdd(&M[(size_t)(B.ao + q) * g.n + A.ao + p], blk[p * nb + q]);
        }
}

// S and T of shell pair k (sa >= sb) of the pair list.
__global__ __launch_bounds__(64) void overlap_kinetic_kernel(OneArgs g) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.npairs) return;
    const int sa = g.pairs[k].x, sb = g.pairs[k].y;
    const OneShell A = g.sh[sa], B = g.sh[sb];
    const int nca = ncart1(A.L), ncb = ncart1(B.L);
    double Sb[kBlk], Tb[kBlk];
    for (int x = 0; x < kBlk; ++x) Sb[x] = Tb[x] = 0.0;
    Herm1<kOneMaxL + 1, kOneMaxL + 3> E[3];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const PrimPairDev pp = prim_pair_dev(g, sa, i, sb, j);
            for (int d = 0; d < 3; ++d) E[d].build(A.L, B.L + 2, pp.p, pp.PA[d], pp.PB[d]);
            const double sq = sqrt(M_PI / pp.p);
            for (int x = 0; x < nca; ++x)
                for (int y = 0; y < ncb; ++y) {
                    int a[3], b[3];
                    cart_lmn(A.L, x, a);
                    cart_lmn(B.L, y, b);
                    const double sx = s1(E[0], sq, a[0], b[0]), sy = s1(E[1], sq, a[1], b[1]),
                                 sz = s1(E[2], sq, a[2], b[2]);
                    const double tx = t1(E[0], sq, pp.b, a[0], b[0]), ty = t1(E[1], sq, pp.b, a[1], b[1]),
                                 tz = t1(E[2], sq, pp.b, a[2], b[2]);
                    Sb[x * ncb + y] += pp.K * sx * sy * sz;
                    Tb[x * ncb + y] += pp.K * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
                }
        }
    double o[kBlk];
    block_out(A.L, B.L, g.pure, Sb, o);
    scatter(g, sa, sb, o, g.S);
    block_out(A.L, B.L, g.pure, Tb, o);
    scatter(g, sa, sb, o, g.T);
}

// The attraction of nucleus c in V for shell pair k (one lane per pair and nucleus, added into
// the zeroed V with atomics).
__global__ __launch_bounds__(64) void nuclear_kernel(OneArgs g) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.npairs * g.natom) return;
    const int k = (int)(t / g.natom), c = (int)(t - (long long)k * g.natom);
    const int sa = g.pairs[k].x, sb = g.pairs[k].y;
    const OneShell A = g.sh[sa], B = g.sh[sb];
    const int nca = ncart1(A.L), ncb = ncart1(B.L);
    double Vb[kBlk];
    for (int x = 0; x < kBlk; ++x) Vb[x] = 0.0;
    Herm1<kOneMaxL + 1, kOneMaxL + 1> E[3];
    double R[kRSize];
    const double zc = g.Z[c];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const PrimPairDev pp = prim_pair_dev(g, sa, i, sb, j);
            for (int d = 0; d < 3; ++d) E[d].build(A.L, B.L, pp.p, pp.PA[d], pp.PB[d]);
            const double pre = 2.0 * M_PI / pp.p * pp.K;
            const double pc[3] = {pp.P[0] - g.xyz[3 * c], pp.P[1] - g.xyz[3 * c + 1], pp.P[2] - g.xyz[3 * c + 2]};
            hermite_r(A.L + B.L, pp.p, pc, g.boys, R);
            for (int x = 0; x < nca; ++x)
                for (int y = 0; y < ncb; ++y) {
                    int a[3], b[3];
                    cart_lmn(A.L, x, a);
                    cart_lmn(B.L, y, b);
                    double acc = 0.0;
                    for (int tt = 0; tt <= a[0] + b[0]; ++tt)
                        for (int u = 0; u <= a[1] + b[1]; ++u)
                            for (int v = 0; v <= a[2] + b[2]; ++v)
                                acc += E[0].get(a[0], b[0], tt) * E[1].get(a[1], b[1], u) * E[2].get(a[2], b[2], v) *
                                       R[ridx(tt, u, v)];
                    Vb[x * ncb + y] -= zc * pre * acc;
                }
        }
    double o[kBlk];
    block_out(A.L, B.L, g.pure, Vb, o);
    scatter_add(g, sa, sb, o, g.V);
}

// One-electron gradient terms of shell pair (sa, sb) and its mirror (sb, sa) (the sum over i, j of
// the gradient formula runs over both orders), added into grad with atomics:
//   2 D_ab d(T + V)_ab/dX - 2 Q_ab dS_ab/dX.
__global__ __launch_bounds__(64) void oneint_grad_kernel(OneArgs g, const double* D, const double* Q, double* grad) {
    // one lane per shell pair (sa >= sb) and part: part natom = overlap and kinetic terms, part
    // c < natom = the attraction of nucleus c. D and Q are symmetric, so the ordered pair (sb, sa)
    // adds what (sa, sb) adds: a pair of two shells counts twice.
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int parts = g.natom + 1;
    if (tid >= (long long)g.npairs * parts) return;
    const int k = (int)(tid / parts), part = (int)(tid - (long long)k * parts);
    const int sa = g.pairs[k].x, sb = g.pairs[k].y;
    const OneShell A = g.sh[sa], B = g.sh[sb];
    const int nca = ncart1(A.L), ncb = ncart1(B.L);
    double WD[kBlk], WQ[kBlk];
    weight_cart(A.L, B.L, g.pure, D, g.n, A.ao, B.ao, WD);
    weight_cart(A.L, B.L, g.pure, Q, g.n, A.ao, B.ao, WQ);
    if (sa != sb)
        for (int x = 0; x < nca * ncb; ++x) {
            WD[x] *= 2.0;
            WQ[x] *= 2.0;
        }
    double gA[3] = {0.0, 0.0, 0.0}, gB[3] = {0.0, 0.0, 0.0};
    Herm1<kOneMaxL + 2, kOneMaxL + 4> E[3];
    double R[kRSize];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const PrimPairDev pp = prim_pair_dev(g, sa, i, sb, j);
            for (int d = 0; d < 3; ++d) E[d].build(A.L + 1, B.L + 3, pp.p, pp.PA[d], pp.PB[d]);
            const double sq = sqrt(M_PI / pp.p);
            // overlap and kinetic: d/dA = 2a |a+1> - a_x |a-1>; d/dB = -d/dA
            if (part == g.natom)
            for (int x = 0; x < nca; ++x)
                for (int y = 0; y < ncb; ++y) {
                    const double wd = WD[x * ncb + y], wq = WQ[x * ncb + y];
                    if (wd == 0.0 && wq == 0.0) continue;
                    int a[3], b[3];
                    cart_lmn(A.L, x, a);
                    cart_lmn(B.L, y, b);
                    double s[3], t[3], ds[3], dt[3];
                    for (int d = 0; d < 3; ++d) {
                        s[d] = s1(E[d], sq, a[d], b[d]);
                        t[d] = t1(E[d], sq, pp.b, a[d], b[d]);
                        ds[d] = 2.0 * pp.a * s1(E[d], sq, a[d] + 1, b[d]) - (a[d] > 0 ? a[d] * s1(E[d], sq, a[d] - 1, b[d]) : 0.0);
                        dt[d] = 2.0 * pp.a * t1(E[d], sq, pp.b, a[d] + 1, b[d]) -
                                (a[d] > 0 ? a[d] * t1(E[d], sq, pp.b, a[d] - 1, b[d]) : 0.0);
                    }
                    for (int d = 0; d < 3; ++d) {
                        const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
                        const double dS = pp.K * ds[d] * s[e1] * s[e2];
                        const double dT = pp.K * (dt[d] * s[e1] * s[e2] + ds[d] * t[e1] * s[e2] + ds[d] * s[e1] * t[e2]);
                        const double v = -2.0 * wq * dS + 2.0 * wd * dT;
                        gA[d] += v;
                        gB[d] -= v;
                    }
                }
            // nuclear attraction: bra and ket shifts, and the operator's centre by translational
            // invariance (dC = -(dA + dB))
            const double pre = 2.0 * M_PI / pp.p * pp.K;
            if (part < g.natom) {
                const int c = part;
                const double pc[3] = {pp.P[0] - g.xyz[3 * c], pp.P[1] - g.xyz[3 * c + 1], pp.P[2] - g.xyz[3 * c + 2]};
                hermite_r(A.L + B.L + 1, pp.p, pc, g.boys, R);
                auto vint = [&](const int* a, const int* b) {
                    if (a[0] < 0 || a[1] < 0 || a[2] < 0 || b[0] < 0 || b[1] < 0 || b[2] < 0) return 0.0;
                    double acc = 0.0;
                    for (int t = 0; t <= a[0] + b[0]; ++t)
                        for (int u = 0; u <= a[1] + b[1]; ++u)
                            for (int v = 0; v <= a[2] + b[2]; ++v)
                                acc += E[0].get(a[0], b[0], t) * E[1].get(a[1], b[1], u) * E[2].get(a[2], b[2], v) *
                                       R[ridx(t, u, v)];
                    return -g.Z[c] * pre * acc;
                };
                double gc[3] = {0.0, 0.0, 0.0};
                for (int x = 0; x < nca; ++x)
                    for (int y = 0; y < ncb; ++y) {
                        const double wd = WD[x * ncb + y];
                        if (wd == 0.0) continue;
                        int a0[3], b0[3];
                        cart_lmn(A.L, x, a0);
                        cart_lmn(B.L, y, b0);
                        for (int d = 0; d < 3; ++d) {
                            int ap[3] = {a0[0], a0[1], a0[2]}, am[3] = {a0[0], a0[1], a0[2]};
                            int bp[3] = {b0[0], b0[1], b0[2]}, bm[3] = {b0[0], b0[1], b0[2]};
                            ap[d] += 1;
                            am[d] -= 1;
                            bp[d] += 1;
                            bm[d] -= 1;
                            const double dA = 2.0 * pp.a * vint(ap, b0) - (a0[d] ? a0[d] * vint(am, b0) : 0.0);
                            const double dB = 2.0 * pp.b * vint(a0, bp) - (b0[d] ? b0[d] * vint(a0, bm) : 0.0);
                            gA[d] += 2.0 * wd * dA;
                            gB[d] += 2.0 * wd * dB;
                            gc[d] -= 2.0 * wd * (dA + dB);
                        }
                    }
                for (int d = 0; d < 3; ++d)
                    if (gc[d] != 0.0) atomicAdd(&grad[3 * c + d], gc[d]);
            }
        }
    for (int d = 0; d < 3; ++d) {
        atomicAdd(&grad[3 * A.atom + d], gA[d]);
        atomicAdd(&grad[3 * B.atom + d], gB[d]);
    }
}

}  // namespace

hipError_t launch_oneint(const OneArgs& g, hipStream_t s) {
    if (g.npairs == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(g.V, 0, (size_t)g.n * g.n * sizeof(double), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(overlap_kinetic_kernel, dim3((g.npairs + 63) / 64), dim3(64), 0, s, g);
    const long long nv = (long long)g.npairs * g.natom;
    hipLaunchKernelGGL(nuclear_kernel, dim3((unsigned)((nv + 63) / 64)), dim3(64), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_oneint_grad(const OneArgs& g, const double* D, const double* Q, double* grad, hipStream_t s) {
    const long long n = (long long)g.npairs * (g.natom + 1);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(oneint_grad_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, g, D, Q, grad);
    return hipGetLastError();
}

}  // namespace hfg

namespace {

int one_fail(hfg_ctx* c, int code, const char* msg) {
    c->err = msg;
    return code;
}

}  // namespace

namespace hfg {

int one_upload(hfg_ctx* c, hfg::OneArgs& g) {
    if (!c->have_basis || !c->have_geom) return one_fail(c, HFG_ERR_ARG, "basis and geometry must be set first");
    if (c->batch || c->mols.size() != 1) return one_fail(c, HFG_ERR_ARG, "one-electron integrals: single-molecule contexts only");
    if ((int)c->Zn.size() != c->natom)
        return one_fail(c, HFG_ERR_ARG, "one-electron integrals: hfg_set_geometry was given no nuclear charges");
    std::vector<hfg::OneShell> sh(c->nshell);
    for (int s = 0; s < c->nshell; ++s)
        sh[s] = hfg::OneShell{c->L[s], c->nprim[s], c->prim_off[s], c->atom[s], c->pure ? c->pure_off[s] : c->cart_off[s]};
    HFG_CHECK(c, hipSetDevice(c->device));
    HFG_CHECK(c, c->d_one_sh.upload(sh, c->stream));
    HFG_CHECK(c, c->d_one_alpha.upload(c->alpha, c->stream));
    HFG_CHECK(c, c->d_one_coef.upload(c->coef, c->stream));
    HFG_CHECK(c, c->d_one_xyz.upload(c->xyz, c->stream));
    HFG_CHECK(c, c->d_one_Z.upload(c->Zn, c->stream));
    g = hfg::OneArgs{};
    g.sh = c->d_one_sh.p;
    g.alpha = c->d_one_alpha.p;
    g.coef = c->d_one_coef.p;
    g.xyz = c->d_one_xyz.p;
    g.Z = c->d_one_Z.p;
    g.boys = c->d_boys.p;
    g.natom = c->natom;
    g.nshell = c->nshell;
    g.n = c->nbf;
    g.pure = c->pure ? 1 : 0;
    return HFG_OK;
}

}  // namespace hfg

extern "C" {

int hfg_oneint(hfg_ctx* c, double* S, double* T, double* V) {
    if (!c) return HFG_ERR_ARG;
    hfg::OneArgs g;
    if (int r = hfg::one_upload(c, g)) return r;
    std::vector<int2> pairs;
    for (int a = 0; a < c->nshell; ++a)
        for (int b = 0; b <= a; ++b) pairs.push_back(make_int2(a, b));
    HFG_CHECK(c, c->d_one_pairs.upload(pairs, c->stream));
    const size_t nn = (size_t)c->nbf * c->nbf;
    HFG_CHECK(c, c->d_one_buf.ensure(3 * nn));
    g.pairs = c->d_one_pairs.p;
    g.npairs = (int)pairs.size();
    g.S = c->d_one_buf.p;
    g.T = g.S + nn;
    g.V = g.T + nn;
    HFG_CHECK(c, hfg::launch_oneint(g, c->stream));
    double* out[3] = {S, T, V};
    for (int k = 0; k < 3; ++k)
        if (out[k]) HFG_CHECK(c, hipMemcpyAsync(out[k], g.S + k * nn, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    return HFG_OK;
}

int hfg_oneint_gradient(hfg_ctx* c, const double* D, const double* Q, double* grad) {
    if (!c) return HFG_ERR_ARG;
    if (!D || !Q || !grad) return one_fail(c, HFG_ERR_ARG, "hfg_oneint_gradient: null argument");
    hfg::OneArgs g;
    if (int r = hfg::one_upload(c, g)) return r;
    std::vector<int2> pairs;
    for (int a = 0; a < c->nshell; ++a)
        for (int b = 0; b <= a; ++b) pairs.push_back(make_int2(a, b));
    HFG_CHECK(c, c->d_one_pairs.upload(pairs, c->stream));
    g.pairs = c->d_one_pairs.p;
    g.npairs = (int)pairs.size();
    const size_t nn = (size_t)c->nbf * c->nbf;
    const int nout = 3 * c->natom;
    HFG_CHECK(c, c->d_one_buf.ensure(2 * nn + nout));
    double* dD = c->d_one_buf.p;
    double* dQ = dD + nn;
    double* dg = dQ + nn;
    HFG_CHECK(c, hipMemcpyAsync(dD, D, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(dQ, Q, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(dg, grad, nout * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, hfg::launch_oneint_grad(g, dD, dQ, dg, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(grad, dg, nout * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    return HFG_OK;
}

}  // extern "C"
