// Two-electron part of the RHF Hessian at a fixed density on the GPU: hfscf's
//   m_hessianC(X1, X2)  =  sum_ijkl D_ij D_kl 2 d2(ij|kl)/dX1 dX2
//   m_hessianEx(X1, X2) = -sum_ijkl D_ij D_kl   d2(ik|jl)/dX1 dX2
// (hfscf: src/hessian/hfscf_hessian_ints.cpp:20-313). hfscf evaluates the second-derivative
// integrals with its chained recursions into six N^4 tensors per direction pair
// (src/integrals/hfscf_oserideriv2qq.cpp:28-749, hfscf_oserideriv2pq.cpp:23-799) and contracts
// them afterwards; here they are contracted with the density as they are made and never stored.
//
// One wavefront per work item: a canonical shell quartet (ab|cd), a >= b, c >= d, (ab) >= (cd),
// and a range of its bra primitive pairs, at most 32 wavefront rounds. Integrals by
// McMurchie-Davidson: per primitive quartet the 1-D Hermite coefficients E^{ij}_t of the bra
// (i <= la + 2, j <= lb + 2) and ket (i <= lc + 2, j <= ld) in LDS, and the Hermite Coulomb
// integrals R_tuv(pq/(p+q), P - Q) up to order la + lb + lc + ld + 2, made level by level. A class
// with at most 32 Cartesian component quartets runs several ket primitive pairs at once, each in a
// slot of lanes with its own ket and R tables. A derivative of a Gaussian with respect to its centre is d/dA_x |a> = 2 alpha |a+1x>
// - a_x |a-1x>, applied once or twice to the centres A and C; the bra's B derivatives come from
// d/dB = d/dP - d/dA with P the bra's Hermite centre (d/dP raises the R index by one), and D from
// translational invariance, d/dD = -(d/dA + d/dB + d/dC). A derivative acts on one direction's
// Hermite coefficients only, so each lane, holding one Cartesian component quartet, forms per
// direction the convolution over the bra and ket Hermite indices of the shifted coefficients (six
// operator combinations) and sums them against R for all 45 coordinate pairs of (A, P, C) x
// (x, y, z).
// The lane sums each pair's integrals over the item's primitive quartets (45 accumulators); at the
// item's end they are weighted by the component quartet's Coulomb and exchange densities, reduced
// over the wavefront, expanded from the 9 x 9 block to the 12 x 12 block of the four centres and
// added to the atoms' rows and columns.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "boys.hpp"
#include "common.hpp"
#include "context.hpp"
#include "coop.hpp"
#include "oneint.hpp"

namespace hfg {

namespace {

constexpr int kHL = 2;                    // largest shell angular momentum
constexpr int kHE = kHL + 3;              // exponents 0 .. L + 2 of a shifted centre
constexpr int kHT = 2 * kHL + 5;          // Hermite index t <= i + j <= 2 L + 4
constexpr int kHN = 4 * kHL + 2;          // R_tuv order <= L_total + 2
constexpr int kHR = (kHN + 1) * (kHN + 2) * (kHN + 3) / 6;
constexpr int kHS = kHN + 1;              // convolution length of one direction
constexpr int kHBlock = 64;               // one wavefront per shell quartet
constexpr int kHPairs = 45;               // coordinate pairs X1 <= X2 over (A, B, C) x (x, y, z)
// LDS for the slots' tables (one slot of the largest class: 3 x 5 x 3 x 7 + 2 x 286 + 11 = 898)
constexpr int kHPool = 1536;

// (t, u, v) of the packed R index ridx(t, u, v) = s(s+1)(s+2)/6 + (s-t)(s-t+1)/2 + v, s = t+u+v.
struct TuvTab {
    unsigned char t[kHR], u[kHR], v[kHR];
};
constexpr TuvTab make_tuv() {
    TuvTab r{};
    int n = 0;
    for (int s = 0; s <= kHN; ++s)
        for (int q = 0; q <= s; ++q)
            for (int v = 0; v <= q; ++v, ++n) {
                r.t[n] = (unsigned char)(s - q);
                r.u[n] = (unsigned char)(q - v);
                r.v[n] = (unsigned char)v;
            }
    return r;
}
__constant__ TuvTab kTuv = make_tuv();

__host__ __device__ constexpr int ridx(int t, int u, int v) {
    const int s = t + u + v;
    return s * (s + 1) * (s + 2) / 6 + (s - t) * (s - t + 1) / 2 + v;
}
__host__ __device__ constexpr int rsize(int n) { return (n + 1) * (n + 2) * (n + 3) / 6; }

// Operators on one direction. Bra codes: 0 none, 1 d/dA, 2 d/dB, 3 d2/dA2, 4 d2/dA dB, 5 d2/dB2;
// ket codes: 0 none, 1 d/dC, 2 d2/dC2. The ten combinations of total order <= 2:
constexpr int kComboBra[10] = {0, 1, 2, 3, 4, 5, 0, 1, 2, 0};
constexpr int kComboKet[10] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 2};
constexpr int kBraOrder[6] = {0, 1, 1, 2, 2, 2};
// the combinations the (A, P, C) coordinates use: none, A, AA, C, AC, CC
constexpr int kUsedCombo[6] = {0, 1, 3, 6, 7, 9};

// Combination per direction of coordinate pair `op` (X1 <= X2, X = 3 centre + direction, in the
// order (0,0) (0,1) .. (0,8) (1,1) .. (8,8)), with the centres taken as A, P, C: P is the bra's
// Hermite centre, whose derivative raises the Hermite index (d/dP_x Lambda_t = Lambda_{t+1}), so
// a P derivative shifts the R index of its direction by one (p[d]) and needs no coefficients of
// its own; d/dB = d/dP - d/dA then gives the B derivatives (the 9 x 9 block is mapped at the end).
struct PairDirs {
    int k[3];
    int p[3];
    int x1, x2;
};
constexpr PairDirs pair_dirs(int op) {
    int X1 = 0, X2 = 0, n = 0;
    for (int a = 0; a < 9; ++a)
        for (int b = a; b < 9; ++b, ++n)
            if (n == op) {
                X1 = a;
                X2 = b;
            }
    int bc[3] = {0, 0, 0}, kc[3] = {0, 0, 0};
    PairDirs r{{0, 0, 0}, {0, 0, 0}, X1, X2};
    const int xs[2] = {X1, X2};
    for (int q = 0; q < 2; ++q) {
        const int c = xs[q] / 3, d = xs[q] % 3;
        if (c == 0)
            bc[d] = bc[d] == 0 ? 1 : 3;
        else if (c == 1)
            r.p[d] += 1;
        else
            kc[d] += 1;
    }
    for (int d = 0; d < 3; ++d) r.k[d] = kc[d] == 0 ? bc[d] : (kc[d] == 2 ? 9 : (bc[d] == 0 ? 6 : 7));
    return r;
}

// pair_dirs of every coordinate pair, for loops over the pairs that are not unrolled.
struct PairTab {
    unsigned char k[kHPairs][3];
    unsigned char p[kHPairs][3];
    unsigned char x1[kHPairs], x2[kHPairs];
};
constexpr PairTab make_pairs() {
    PairTab t{};
    for (int op = 0; op < kHPairs; ++op) {
        const PairDirs pd = pair_dirs(op);
        for (int d = 0; d < 3; ++d) {
            t.k[op][d] = (unsigned char)pd.k[d];
            t.p[op][d] = (unsigned char)pd.p[d];
        }
        t.x1[op] = (unsigned char)pd.x1;
        t.x2[op] = (unsigned char)pd.x2;
    }
    return t;
}
__constant__ PairTab kPairs = make_pairs();

// How a wavefront lays out a class: ncpad lanes per slot (the component quartets rounded up to a
// power of two, 64 above 32), nslot slots, each with its own ket tables (ksj, kst: strides of the
// ket E^{ij}_t; kdir doubles per direction) and R tables (rn entries).
struct SlotShape {
    int ncomp, ncpad, nslot, ksj, kst, kdir, rn, slot_size;
};
__host__ __device__ inline SlotShape slot_shape(int la, int lb, int lc, int ld) {
    SlotShape s;
    s.ncomp = (la + 1) * (la + 2) / 2 * ((lb + 1) * (lb + 2) / 2) * ((lc + 1) * (lc + 2) / 2) * ((ld + 1) * (ld + 2) / 2);
    s.ncpad = s.ncomp > 32 ? 64 : s.ncomp > 16 ? 32 : s.ncomp > 8 ? 16 : s.ncomp > 4 ? 8 : s.ncomp > 2 ? 4 : s.ncomp;
    s.ksj = ld + 1;
    s.kst = lc + ld + 3;
    s.kdir = (lc + 3) * s.ksj * s.kst;
    s.rn = rsize(la + lb + lc + ld + 2);
    s.slot_size = 3 * s.kdir + 2 * s.rn + la + lb + lc + ld + 3;
    s.nslot = kHBlock / s.ncpad < kHPool / s.slot_size ? kHBlock / s.ncpad : kHPool / s.slot_size;
    return s;
}

// A work item: canonical shell-pair indices P >= Q and a range [b0, b1) of the quartet's primitive
// work (bra primitive pairs for the generic kernel, flattened primitive quartets bra-major for the
// per-lane kernels), so that no item holds more than a few hundred primitive quartets.
struct HItem {
    int P, Q, b0, b1;
};

struct HessArgs {
    const OneShell* sh;
    const double* alpha;
    const double* coef;
    const double* xyz;
    const double* boys;
    const int2* pairs;    // canonical shell pairs (a, b), a >= b
    const int* cart_off;  // first Cartesian function of each shell
    int ncart, natom;
    const double* Dc;     // ncart x ncart density in the Cartesian basis
    double prim_cut;      // primitive quartets with |prefactor| below it are skipped
    double* HC;           // (3 natom)^2 Coulomb part (zeroed before the launch)
    double* HX;           // (3 natom)^2 exchange part
};

HFG_INLINE void cart_exps(int L, int k, int* e) {
    int n = 0;
    for (int i = 0; i <= L; ++i)
        for (int j = 0; j <= i; ++j, ++n)
            if (n == k) {
                e[0] = L - i;
                e[1] = i - j;
                e[2] = j;
            }
}

// E^{ij}_t of one direction at E[(i * sj + j) * st + t]; zero outside 0 <= t <= i + j.
HFG_INLINE double eget(const double* E, int sj, int st, int i, int j, int t) {
    return (i < 0 || j < 0 || t < 0 || t > i + j) ? 0.0 : E[(i * sj + j) * st + t];
}

HFG_INLINE void herm_fill(double* E, int sj, int st, int imax, int jmax, double o2p, double xpa, double xpb) {
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j)
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i == 0 && j == 0)
                    v = 1.0;
                else if (i > 0)
                    v = o2p * eget(E, sj, st, i - 1, j, t - 1) + xpa * eget(E, sj, st, i - 1, j, t) +
                        (t + 1) * eget(E, sj, st, i - 1, j, t + 1);
                else
                    v = o2p * eget(E, sj, st, i, j - 1, t - 1) + xpb * eget(E, sj, st, i, j - 1, t) +
                        (t + 1) * eget(E, sj, st, i, j - 1, t + 1);
                E[(i * sj + j) * st + t] = v;
            }
}

// Shifted bra coefficient of code `code` (za = 2 alpha, zb = 2 beta).
HFG_INLINE double bra_op(const double* E, int code, int a, int b, int t, double za, double zb) {
    constexpr int sj = kHE;
    switch (code) {
        case 0: return eget(E, sj, kHT, a, b, t);
        case 1: return za * eget(E, sj, kHT, a + 1, b, t) - a * eget(E, sj, kHT, a - 1, b, t);
        case 2: return zb * eget(E, sj, kHT, a, b + 1, t) - b * eget(E, sj, kHT, a, b - 1, t);
        case 3:
            return za * za * eget(E, sj, kHT, a + 2, b, t) - za * (2 * a + 1) * eget(E, sj, kHT, a, b, t) +
                   a * (a - 1) * eget(E, sj, kHT, a - 2, b, t);
        case 4:
            return za * zb * eget(E, sj, kHT, a + 1, b + 1, t) - za * b * eget(E, sj, kHT, a + 1, b - 1, t) -
                   zb * a * eget(E, sj, kHT, a - 1, b + 1, t) + a * b * eget(E, sj, kHT, a - 1, b - 1, t);
        default:
            return zb * zb * eget(E, sj, kHT, a, b + 2, t) - zb * (2 * b + 1) * eget(E, sj, kHT, a, b, t) +
                   b * (b - 1) * eget(E, sj, kHT, a, b - 2, t);
    }
}

// Shifted ket coefficient (zc = 2 gamma; the ket table's strides sj, st).
HFG_INLINE double ket_op(const double* E, int sj, int st, int code, int c, int d, int t, double zc) {
    switch (code) {
        case 0: return eget(E, sj, st, c, d, t);
        case 1: return zc * eget(E, sj, st, c + 1, d, t) - c * eget(E, sj, st, c - 1, d, t);
        default:
            return zc * zc * eget(E, sj, st, c + 2, d, t) - zc * (2 * c + 1) * eget(E, sj, st, c, d, t) +
                   c * (c - 1) * eget(E, sj, st, c - 2, d, t);
    }
}

HFG_INLINE void boys_upto(int N, double T, const double* tab, double* F) {
    switch (N) {
        case 2: boys_eval<2, double>(T, tab, 1.0, F); break;
        case 3: boys_eval<3, double>(T, tab, 1.0, F); break;
        case 4: boys_eval<4, double>(T, tab, 1.0, F); break;
        case 5: boys_eval<5, double>(T, tab, 1.0, F); break;
        case 6: boys_eval<6, double>(T, tab, 1.0, F); break;
        case 7: boys_eval<7, double>(T, tab, 1.0, F); break;
        case 8: boys_eval<8, double>(T, tab, 1.0, F); break;
        case 9: boys_eval<9, double>(T, tab, 1.0, F); break;
        default: boys_eval<kHN, double>(T, tab, 1.0, F); break;
    }
}

HFG_INLINE double T_of(double aR, const double* PQ) { return aR * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]); }

HFG_INLINE double wave_sum(double v) {
    for (int o = kHBlock / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kHBlock);
    return v;
}

// Atom entries of HC and HX from the item's symmetric 9 x 9 blocks M over (A, B, C) x (x, y, z):
// the 12 x 12 block of the four centres (the D rows and columns -(A + B + C)) added with atomics.
HFG_INLINE void finish_item(const HessArgs& g, const int atoms[4], double (*M)[81]) {
    const int n3 = 3 * g.natom;
    for (int y = threadIdx.x; y < 144; y += kHBlock) {
        const int y1 = y / 12, y2 = y % 12;
        double vc = 0.0, vx = 0.0;
        for (int c1 = 0; c1 < 3; ++c1) {
            if (y1 < 9 && c1 != y1 / 3) continue;
            const int x1 = 3 * c1 + y1 % 3;
            const double t1 = y1 < 9 ? 1.0 : -1.0;
            for (int c2 = 0; c2 < 3; ++c2) {
                if (y2 < 9 && c2 != y2 / 3) continue;
                const int x2 = 3 * c2 + y2 % 3;
                const double t = y2 < 9 ? t1 : -t1;
                vc += t * M[0][x1 * 9 + x2];
                vx += t * M[1][x1 * 9 + x2];
            }
        }
        const size_t r = (size_t)atoms[y1 / 3] * 3 + y1 % 3, col = (size_t)atoms[y2 / 3] * 3 + y2 % 3;
        unsafeAtomicAdd(&g.HC[r * n3 + col], vc);
        unsafeAtomicAdd(&g.HX[r * n3 + col], vx);
    }
}

__global__ __launch_bounds__(kHBlock) __attribute__((amdgpu_num_vgpr(128))) void hess2e_kernel(HessArgs g, const HItem* items,
                                                                                          int nitems) {
    __shared__ double Eb[3 * kHE * kHE * kHT];
    __shared__ double pool[kHPool];  // per slot: ket E^{ij}_t (3 directions), R ping-pong, scaled F_m
    __shared__ double M[2][81];  // the item's 9 x 9 blocks over (A, P, C), then (A, B, C), x (x, y, z)
    const int lane = threadIdx.x;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const HItem w = items[it];
        const int2 pab = g.pairs[w.P], pcd = g.pairs[w.Q];
        const int s4[4] = {pab.x, pab.y, pcd.x, pcd.y};
        const OneShell A = g.sh[s4[0]], B = g.sh[s4[1]], C = g.sh[s4[2]], D = g.sh[s4[3]];
        const double* ra = g.xyz + 3 * A.atom;
        const double* rb = g.xyz + 3 * B.atom;
        const double* rc = g.xyz + 3 * C.atom;
        const double* rd = g.xyz + 3 * D.atom;
        double deg = 8.0;
        if (pab.x == pab.y) deg *= 0.5;
        if (pcd.x == pcd.y) deg *= 0.5;
        if (w.P == w.Q) deg *= 0.5;
        const int nb = (B.L + 1) * (B.L + 2) / 2;
        const int nc = (C.L + 1) * (C.L + 2) / 2, nd = (D.L + 1) * (D.L + 2) / 2;
        const SlotShape shp = slot_shape(A.L, B.L, C.L, D.L);
        const int ncomp = shp.ncomp;
        const int N = A.L + B.L + C.L + D.L + 2;
        const int RN = shp.rn;
        const int oa = g.cart_off[s4[0]], ob = g.cart_off[s4[1]], oc = g.cart_off[s4[2]], od = g.cart_off[s4[3]];
        double ab2 = 0.0, cd2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            ab2 += (ra[d] - rb[d]) * (ra[d] - rb[d]);
            cd2 += (rc[d] - rd[d]) * (rc[d] - rd[d]);
        }
        // slots: groups of ncpad lanes, each on its own ket primitive pair with its own tables
        const int ncpad = shp.ncpad, ksj = shp.ksj, kst = shp.kst, kdir = shp.kdir;
        const int slot_size = shp.slot_size, nslot = shp.nslot;
        const int slot = lane / ncpad, sl = lane - slot * ncpad;
        double* Ek = pool + (slot < nslot ? slot : 0) * slot_size;
        double* Rb0 = Ek + 3 * kdir;
        double* Rb1 = Rb0 + RN;
        double* Fs = Rb1 + RN;
        const int nket = C.np * D.np;
        for (int k = lane; k < 162; k += kHBlock) M[k / 81][k % 81] = 0.0;
        // chunks of 64 component quartets (one unless the class has more), each over every
        // primitive quartet of the item, with one component quartet per lane
        for (int cbase = 0; cbase < ncomp; cbase += kHBlock) {
            const int cq = cbase + sl;
            const bool has_comp = cq < ncomp;
            const int ia = cq / (nb * nc * nd), ib = (cq / (nc * nd)) % nb, ic = (cq / nd) % nc, id = cq % nd;
            int ea[3] = {0, 0, 0}, eb[3] = {0, 0, 0}, ec[3] = {0, 0, 0}, ed[3] = {0, 0, 0};
            if (has_comp) {
                cart_exps(A.L, ia, ea);
                cart_exps(B.L, ib, eb);
                cart_exps(C.L, ic, ec);
                cart_exps(D.L, id, ed);
            }
            double acc[kHPairs];  // per coordinate pair: sum of pref x integral over the item
            for (int op = 0; op < kHPairs; ++op) acc[op] = 0.0;
            for (int bp = w.b0; bp < w.b1; ++bp) {
                const int i = bp / B.np, j = bp % B.np;
                const double al = g.alpha[A.po + i], be = g.alpha[B.po + j], p = al + be;
                const double Kab = exp(-al * be / p * ab2) * g.coef[A.po + i] * g.coef[B.po + j];
                double Pc[3];
                for (int d = 0; d < 3; ++d) Pc[d] = (al * ra[d] + be * rb[d]) / p;
                // the previous use of Eb ended at a barrier
                if (lane < 3)
                    herm_fill(Eb + lane * kHE * kHE * kHT, kHE, kHT, A.L + 2, B.L + 2, 0.5 / p, Pc[lane] - ra[lane],
                              Pc[lane] - rb[lane]);
                for (int kbase = 0; kbase < nket; kbase += nslot) {
                    const int kp = kbase + slot;
                    bool active = slot < nslot && kp < nket;
                    double ga = 1.0, pref = 0.0, aR = 1.0, PQ[3] = {0.0, 0.0, 0.0};
                    if (active) {
                        const int k = kp / D.np, l = kp % D.np;
                        ga = g.alpha[C.po + k];
                        const double de = g.alpha[D.po + l], qx = ga + de;
                        const double Kcd = exp(-ga * de / qx * cd2) * g.coef[C.po + k] * g.coef[D.po + l];
                        pref = 34.986836655249725693 * Kab * Kcd / (p * qx * sqrt(p + qx));  // 2 pi^2.5
                        active = fabs(pref) >= g.prim_cut;
                        double Qc[3];
                        for (int d = 0; d < 3; ++d) {
                            Qc[d] = (ga * rc[d] + de * rd[d]) / qx;
                            PQ[d] = Pc[d] - Qc[d];
                        }
                        aR = p * qx / (p + qx);
                        if (active) {
                            for (int d = sl; d < 3; d += ncpad)
                                herm_fill(Ek + d * kdir, ksj, kst, C.L + 2, D.L, 0.5 / qx, Qc[d] - rc[d], Qc[d] - rd[d]);
                            if (sl == 0) {
                                double F[kHN + 1];
                                boys_upto(N, T_of(aR, PQ), g.boys, F);
                                double f = 1.0;
                                for (int m = 0; m <= N; ++m) {
                                    Fs[m] = f * F[m];
                                    f *= -2.0 * aR;
                                }
                            }
                        }
                    }
                    __syncthreads();
                    // R^(m)_tuv from order N down to 0 (level m in Rb0 for even m, Rb1 for odd m)
                    if (active && sl == 0) ((N & 1) ? Rb1 : Rb0)[0] = Fs[N];
                    __syncthreads();
                    for (int m = N - 1; m >= 0; --m) {
                        if (active) {
                            const double* src = ((m + 1) & 1) ? Rb1 : Rb0;
                            double* dst = (m & 1) ? Rb1 : Rb0;
                            const int cnt = rsize(N - m);
                            for (int x = sl; x < cnt; x += ncpad) {
                                const int t = kTuv.t[x], u = kTuv.u[x], v = kTuv.v[x];
                                double val;
                                if (x == 0)
                                    val = Fs[m];
                                else if (t > 0)
                                    val = (t > 1 ? (t - 1) * src[ridx(t - 2, u, v)] : 0.0) + PQ[0] * src[ridx(t - 1, u, v)];
                                else if (u > 0)
                                    val = (u > 1 ? (u - 1) * src[ridx(t, u - 2, v)] : 0.0) + PQ[1] * src[ridx(t, u - 1, v)];
                                else
                                    val = (v > 1 ? (v - 1) * src[ridx(t, u, v - 2)] : 0.0) + PQ[2] * src[ridx(t, u, v - 1)];
                                dst[x] = val;
                            }
                        }
                        __syncthreads();
                    }
                    if (active && has_comp) {
                        const double* R = Rb0;
                        const double za = 2.0 * al, zb = 2.0 * be, zc = 2.0 * ga;
                        // per direction and combination: sum over t + tau = S of bra x (-1)^tau ket
                        double cv[3][10][kHS];
                        int len[3][10];
                        for (int d = 0; d < 3; ++d) {
                            const double* Ebd = Eb + d * kHE * kHE * kHT;
                            const double* Ekd = Ek + d * kdir;
                            for (int u = 0; u < 6; ++u) {
                                const int cb = kUsedCombo[u];
                                const int bcode = kComboBra[cb], kcode = kComboKet[cb];
                                const int nt = ea[d] + eb[d] + kBraOrder[bcode] + 1;
                                const int nk = ec[d] + ed[d] + kcode + 1;
                                double bv[kHT], kv[kHT];
                                for (int t = 0; t < nt; ++t) bv[t] = bra_op(Ebd, bcode, ea[d], eb[d], t, za, zb);
                                for (int t = 0; t < nk; ++t)
                                    kv[t] = ((t & 1) ? -1.0 : 1.0) * ket_op(Ekd, ksj, kst, kcode, ec[d], ed[d], t, zc);
                                len[d][cb] = nt + nk - 1;
                                for (int S = 0; S < nt + nk - 1; ++S) {
                                    double sacc = 0.0;
                                    for (int t = (S - nk + 1 > 0 ? S - nk + 1 : 0); t < nt && t <= S; ++t)
                                        sacc += bv[t] * kv[S - t];
                                    cv[d][cb][S] = sacc;
                                }
                            }
                        }
#pragma unroll 1
                        for (int op = 0; op < kHPairs; ++op) {
                            const int kx = kPairs.k[op][0], ky = kPairs.k[op][1], kz = kPairs.k[op][2];
                            const int px = kPairs.p[op][0], py = kPairs.p[op][1], pz = kPairs.p[op][2];
                            const double* cx = cv[0][kx];
                            const double* cy = cv[1][ky];
                            const double* cz = cv[2][kz];
                            const int lx = len[0][kx], ly = len[1][ky], lz = len[2][kz];
                            double czr[kHS];  // the innermost factor in registers
#pragma unroll
                            for (int V = 0; V < kHS; ++V) czr[V] = V < lz ? cz[V] : 0.0;
                            double I = 0.0;
                            for (int S = 0; S < lx; ++S) {
                                double sy = 0.0;
                                for (int U = 0; U < ly; ++U) {
                                    // ridx(t, u, v), t = S + px, u = U + py, v = V + pz, stepped in V:
                                    // the step is (s+1)(s+2)/2 + (u+v) + 2 at s = t + u + v, and grows
                                    // by s + 3 per step
                                    const int t0 = S + px, u0 = U + py, s0 = t0 + u0 + pz;
                                    int idx = ridx(t0, u0, pz), ds = (s0 + 1) * (s0 + 2) / 2 + u0 + pz + 2;
                                    double sz = 0.0;
#pragma unroll
                                    for (int V = 0; V < kHS; ++V) {
                                        if (V < lz) {
                                            sz += czr[V] * R[idx];
                                            idx += ds;
                                            ds += s0 + V + 3;  // s = s0 + V at this step
                                        }
                                    }
                                    sy += cy[U] * sz;
                                }
                                I += cx[S] * sy;
                            }
                            acc[op] += pref * I;
                        }
                    }
                    __syncthreads();  // the slots' tables are rewritten by the next round
                }
            }
            // densities of the lane's component quartet, the wavefront's sums into M
            double gC = 0.0, gX = 0.0;
            if (has_comp) {
                const int fa = oa + ia, fb = ob + ib, fc = oc + ic, fd = od + id;
                const size_t n = (size_t)g.ncart;
                gC = deg * 2.0 * g.Dc[fa * n + fb] * g.Dc[fc * n + fd];
                gX = -deg * 0.5 * (g.Dc[fa * n + fc] * g.Dc[fb * n + fd] + g.Dc[fa * n + fd] * g.Dc[fb * n + fc]);
            }
#pragma unroll 1
            for (int op = 0; op < kHPairs; ++op) {
                const int x1 = kPairs.x1[op], x2 = kPairs.x2[op];
                const double c = wave_sum(gC * acc[op]), x = wave_sum(gX * acc[op]);
                if (lane == 0) {
                    M[0][x1 * 9 + x2] += c;
                    M[1][x1 * 9 + x2] += x;
                    if (x1 != x2) {
                        M[0][x2 * 9 + x1] += c;
                        M[1][x2 * 9 + x1] += x;
                    }
                }
            }
        }
        __syncthreads();
        // (A, P, C) -> (A, B, C): d/dB = d/dP - d/dA in each direction (in place, through registers)
        double mv[(162 + kHBlock - 1) / kHBlock];
        for (int r = 0, e = lane; e < 162; ++r, e += kHBlock) {
            const int m = e / 81, x1 = (e % 81) / 9, x2 = e % 9;
            double v = 0.0;
            for (int q1 = 0; q1 < 2; ++q1) {
                // row x1 of the map: A -> A; B -> P - A; C -> C
                const int c1 = x1 / 3;
                if (q1 == 1 && c1 != 1) continue;
                const int y1 = c1 == 1 ? (q1 == 0 ? x1 : x1 - 3) : x1;
                const double w1 = (c1 == 1 && q1 == 1) ? -1.0 : 1.0;
                for (int q2 = 0; q2 < 2; ++q2) {
                    const int c2 = x2 / 3;
                    if (q2 == 1 && c2 != 1) continue;
                    const int y2 = c2 == 1 ? (q2 == 0 ? x2 : x2 - 3) : x2;
                    const double w2 = (c2 == 1 && q2 == 1) ? -1.0 : 1.0;
                    v += w1 * w2 * M[m][y1 * 9 + y2];
                }
            }
            mv[r] = v;
        }
        __syncthreads();
        for (int r = 0, e = lane; e < 162; ++r, e += kHBlock) M[e / 81][e % 81] = mv[r];
        __syncthreads();
        const int atoms[4] = {A.atom, B.atom, C.atom, D.atom};
        finish_item(g, atoms, M);
        __syncthreads();  // M is rewritten by the next item
    }
}

}  // namespace

}  // namespace hfg

extern "C" {

int hfg_hessian_2e(hfg_ctx* c, const double* D, double* HC, double* HX) {
    if (!c) return HFG_ERR_ARG;
    if (!D || !HC || !HX) {
        c->err = "hfg_hessian_2e: null argument";
        return HFG_ERR_ARG;
    }
    hfg::OneArgs one;
    if (int r = hfg::one_upload(c, one)) return r;
    for (int s = 0; s < c->nshell; ++s)
        if (c->L[s] > hfg::kHL) {
            c->err = "hfg_hessian_2e: shells up to d";
            return HFG_ERR_ARG;
        }
    // the density in the Cartesian basis: Dc = T D T^T with T the output-basis -> Cartesian map
    const int n = c->nbf, nc = c->ncart;
    std::vector<double> Dc((size_t)nc * nc, 0.0);
    auto terms = [&](int s, int k, int* cidx, double* w) {  // output function k of shell s
        if (!c->pure || c->L[s] == 0) {
            cidx[0] = k;
            w[0] = 1.0;
            return 1;
        }
        const hfg::SphTerms t = hfg::sph_terms(c->L[s], k);
        for (int x = 0; x < t.n; ++x) {
            cidx[x] = t.c[x];
            w[x] = t.w[x];
        }
        return t.n;
    };
    const std::vector<int>& out_off = c->pure ? c->pure_off : c->cart_off;
    for (int sa = 0; sa < c->nshell; ++sa)
        for (int sb = 0; sb < c->nshell; ++sb) {
            const int La = c->L[sa], Lb = c->L[sb];
            const int noa = c->pure ? 2 * La + 1 : (La + 1) * (La + 2) / 2;
            const int nob = c->pure ? 2 * Lb + 1 : (Lb + 1) * (Lb + 2) / 2;
            for (int p = 0; p < noa; ++p) {
                int ca[3];
                double wa[3];
                const int ka = terms(sa, p, ca, wa);
                for (int q = 0; q < nob; ++q) {
                    int cb[3];
                    double wb[3];
                    const int kb = terms(sb, q, cb, wb);
                    const double v = D[(size_t)(out_off[sa] + p) * n + out_off[sb] + q];
                    for (int x = 0; x < ka; ++x)
                        for (int y = 0; y < kb; ++y)
                            Dc[(size_t)(c->cart_off[sa] + ca[x]) * nc + c->cart_off[sb] + cb[y]] += wa[x] * wb[y] * v;
                }
            }
        }
    std::vector<int2> pairs;
    for (int a = 0; a < c->nshell; ++a)
        for (int b = 0; b <= a; ++b) pairs.push_back(make_int2(a, b));
    const int n3 = 3 * c->natom;
    const size_t nh = (size_t)n3 * n3;
    HFG_CHECK(c, c->d_one_pairs.upload(pairs, c->stream));
    HFG_CHECK(c, c->d_hess_cart.upload(c->cart_off, c->stream));
    HFG_CHECK(c, c->d_one_buf.ensure(Dc.size() + 2 * nh));
    double* dDc = c->d_one_buf.p;
    double* dH = dDc + Dc.size();
    HFG_CHECK(c, hipMemcpyAsync(dDc, Dc.data(), Dc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HFG_CHECK(c, hipMemsetAsync(dH, 0, 2 * nh * sizeof(double), c->stream));
    hfg::HessArgs g{};
    g.sh = one.sh;
    g.alpha = one.alpha;
    g.coef = one.coef;
    g.xyz = one.xyz;
    g.boys = one.boys;
    g.pairs = c->d_one_pairs.p;
    g.cart_off = c->d_hess_cart.p;
    g.ncart = nc;
    g.natom = c->natom;
    g.Dc = dDc;
    g.prim_cut = c->opt.integral_tol / 1e4;
    g.HC = dH;
    g.HX = dH + nh;
    // work items of at most kItemRounds wavefront rounds (component chunks x rounds of ket primitive
    // pairs over the slots x bra primitive pairs, as the kernel lays them out), in whole bra
    // primitive pairs, grouped by the order N = la + lb + lc + ld + 2
    constexpr int kItemRounds = 32;
    std::vector<std::vector<hfg::HItem>> by_n(hfg::kHN + 1);
    const int np = (int)pairs.size();
    for (int P = 0; P < np; ++P)
        for (int Q = 0; Q <= P; ++Q) {
            const int a = pairs[P].x, b = pairs[P].y, cc = pairs[Q].x, d = pairs[Q].y;
            const int N = c->L[a] + c->L[b] + c->L[cc] + c->L[d] + 2;
            const int nbra = c->nprim[a] * c->nprim[b], nket = c->nprim[cc] * c->nprim[d];
            const hfg::SlotShape sh = hfg::slot_shape(c->L[a], c->L[b], c->L[cc], c->L[d]);
            const int chunks = (sh.ncomp + hfg::kHBlock - 1) / hfg::kHBlock;
            const int rounds_per_bra = chunks * ((nket + sh.nslot - 1) / sh.nslot);
            const int per = std::max(1, kItemRounds / rounds_per_bra);
            for (int b0 = 0; b0 < nbra; b0 += per) by_n[N].push_back(hfg::HItem{P, Q, b0, std::min(nbra, b0 + per)});
        }
    // heaviest classes first, so that the launch does not end on a few long items
    std::vector<int4> packed;
    for (int N = hfg::kHN; N >= 0; --N)
        for (const hfg::HItem& it : by_n[N]) packed.push_back(make_int4(it.P, it.Q, it.b0, it.b1));
    HFG_CHECK(c, c->d_hess_items.upload(packed, c->stream));
    const hfg::HItem* d_items = reinterpret_cast<const hfg::HItem*>(c->d_hess_items.p);
    static_assert(sizeof(hfg::HItem) == sizeof(int4), "work items travel as int4");
    // the kernel spills to scratch: it runs on the device's shared scratch stream, after the uploads
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    if (!packed.empty())
        hipLaunchKernelGGL(hfg::hess2e_kernel, dim3((unsigned)std::min<size_t>(packed.size(), 1 << 16)), dim3(hfg::kHBlock),
                           0, c->heavy, g, d_items, (int)packed.size());
    HFG_CHECK(c, hipGetLastError());
    HFG_CHECK(c, hipStreamSynchronize(c->heavy));
    HFG_CHECK(c, hipMemcpyAsync(HC, g.HC, nh * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipMemcpyAsync(HX, g.HX, nh * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HFG_CHECK(c, hipStreamSynchronize(c->stream));
    return HFG_OK;
}

}  // extern "C"
