"""ORACLE — test infrastructure only. Independent McMurchie-Davidson ERIs in numpy.

Textbook scheme (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory",
sections 9.5 and 9.9): Hermite expansion coefficients E^{ij}_t per Cartesian direction and
Hermite Coulomb integrals R_{tuv} built from the Boys function, which here comes from scipy
(regularised lower incomplete gamma). It shares no code with the oracle's chained 1-D
Obara-Saika recursions or with the GPU's Head-Gordon-Pople kernels, so agreement pins all
three. Slow (pure Python loops): small cases only.
"""
import math

import numpy as np
from scipy.special import gamma, gammainc


def boys(m, T):
    """F_m(T) from the regularised incomplete gamma function, series for tiny T."""
    if T < 1e-6:
        return sum((-T) ** k / (math.factorial(k) * (2 * m + 2 * k + 1)) for k in range(6))
    a = m + 0.5
    return gamma(a) * gammainc(a, T) / (2.0 * T ** a)


def comps(l):
    return [(l - i, i - j, j) for i in range(l + 1) for j in range(i + 1)]


def hermite_e(i, j, a, b, xa, xb):
    """E^{ij}_t, t = 0..i+j, for exponents a, b at centres xa, xb (1-D)."""
    p = a + b
    mu = a * b / p
    X = xa - xb
    E = np.zeros((i + 1, j + 1, i + j + 2))
    E[0, 0, 0] = math.exp(-mu * X * X)
    xpa = -b / p * X
    xpb = a / p * X
    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = (xpa * E[ii - 1, jj, t] + (t + 1) * E[ii - 1, jj, t + 1])
                    if t > 0:
                        v += E[ii - 1, jj, t - 1] / (2 * p)
                else:
                    v = (xpb * E[ii, jj - 1, t] + (t + 1) * E[ii, jj - 1, t + 1])
                    if t > 0:
                        v += E[ii, jj - 1, t - 1] / (2 * p)
                E[ii, jj, t] = v
    return E[:, :, : i + j + 1]


def hermite_r(N, p, PC):
    """R_{tuv} (t+u+v <= N) for exponent p and vector PC."""
    T = p * float(np.dot(PC, PC))
    R = np.zeros((N + 1, N + 1, N + 1, N + 1))  # [n][t][u][v]
    for n in range(N + 1):
        R[n, 0, 0, 0] = (-2 * p) ** n * boys(n, T)
    for s in range(1, N + 1):
        for n in range(N - s + 1):
            for t in range(s + 1):
                for u in range(s - t + 1):
                    v = s - t - u
                    if t > 0:
                        val = PC[0] * R[n + 1, t - 1, u, v] + ((t - 1) * R[n + 1, t - 2, u, v] if t > 1 else 0)
                    elif u > 0:
                        val = PC[1] * R[n + 1, t, u - 1, v] + ((u - 1) * R[n + 1, t, u - 2, v] if u > 1 else 0)
                    else:
                        val = PC[2] * R[n + 1, t, u, v - 1] + ((v - 1) * R[n + 1, t, u, v - 2] if v > 1 else 0)
                    R[n, t, u, v] = val
    return R[0]


def primitive_eri(la, A, a, lb, B, b, lc, C, c, ld, D, d):
    """All Cartesian components of one primitive quartet, shape (ncart^4)."""
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    alpha = p * q / (p + q)
    Eab = [hermite_e(la, lb, a, b, A[k], B[k]) for k in range(3)]
    Ecd = [hermite_e(lc, ld, c, d, C[k], D[k]) for k in range(3)]
    R = hermite_r(la + lb + lc + ld, alpha, P - Q)
    pre = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    ca, cb, cc, cd = comps(la), comps(lb), comps(lc), comps(ld)
    out = np.zeros((len(ca), len(cb), len(cc), len(cd)))
    for i1, e1 in enumerate(ca):
        for i2, e2 in enumerate(cb):
            # bra Hermite coefficients as a tensor over (t, u, v)
            Ex = Eab[0][e1[0], e2[0]]
            Ey = Eab[1][e1[1], e2[1]]
            Ez = Eab[2][e1[2], e2[2]]
            for i3, e3 in enumerate(cc):
                for i4, e4 in enumerate(cd):
                    Fx = Ecd[0][e3[0], e4[0]]
                    Fy = Ecd[1][e3[1], e4[1]]
                    Fz = Ecd[2][e3[2], e4[2]]
                    s = 0.0
                    for t in range(len(Ex)):
                        for u in range(len(Ey)):
                            for v in range(len(Ez)):
                                ebra = Ex[t] * Ey[u] * Ez[v]
                                if ebra == 0.0:
                                    continue
                                for tau in range(len(Fx)):
                                    for nu in range(len(Fy)):
                                        for phi in range(len(Fz)):
                                            s += ebra * Fx[tau] * Fy[nu] * Fz[phi] * (-1) ** (tau + nu + phi) * \
                                                R[t + tau, u + nu, v + phi]
                    out[i1, i2, i3, i4] = pre * s
    return out


def contracted_eri(shells, xyz, s1, s2, s3, s4, prim_cut=0.0):
    """Contracted Cartesian block using the normalised coefficients in shells['coef'].

    prim_cut > 0 drops primitive quartets whose prefactor
    |c_i c_j c_k c_l e^{-ab/p AB^2} e^{-cd/q CD^2} 2 pi^{5/2} / (p q sqrt(p+q))| is below it."""
    po = np.concatenate([[0], np.cumsum(shells["nprim"])])
    L = shells["L"]
    cen = [np.asarray(xyz[shells["atom"][s]], dtype=float) for s in (s1, s2, s3, s4)]
    al = shells["alpha"]
    ab2 = float(np.sum((cen[0] - cen[1]) ** 2))
    cd2 = float(np.sum((cen[2] - cen[3]) ** 2))
    res = np.zeros([len(comps(L[s])) for s in (s1, s2, s3, s4)])
    for i in range(po[s1], po[s1 + 1]):
        for j in range(po[s2], po[s2 + 1]):
            for k in range(po[s3], po[s3 + 1]):
                for l in range(po[s4], po[s4 + 1]):
                    w = shells["coef"][i] * shells["coef"][j] * shells["coef"][k] * shells["coef"][l]
                    if prim_cut > 0.0:
                        p, q = al[i] + al[j], al[k] + al[l]
                        pf = w * math.exp(-al[i] * al[j] / p * ab2 - al[k] * al[l] / q * cd2) * 2 * math.pi ** 2.5 / (
                            p * q * math.sqrt(p + q))
                        if abs(pf) < prim_cut:
                            continue
                    blk = w * primitive_eri(L[s1], cen[0], shells["alpha"][i], L[s2], cen[1], shells["alpha"][j],
                                            L[s3], cen[2], shells["alpha"][k], L[s4], cen[3], shells["alpha"][l])
                    res = res + blk
    return res
