"""ORACLE — test infrastructure only. One-electron integrals, their nuclear derivatives, the
nuclear repulsion and the RHF gradient assembly, in numpy.

Restates what hfscf computes on the host around the two-electron path:
  * S, T, V and H = T + V       src/scfsolver/hfscf_hf_hamiltonian.cpp:14-38,
                                src/integrals/hfscf_osoverlap.cpp:10, hfscf_osnuclear.cpp:17-83
  * dS, dT, dV per atom         src/integrals/hfscf_osoverlapderiv1.cpp:8,
                                hfscf_oskineticderiv1.cpp:8, hfscf_osnuclearderiv1.cpp:15
  * E_NN and its gradient       src/molecule/hfscf_molecule.cpp:338-352
  * g = sum D.dGdX + 2 D.(dT + dV) - 2 (D F D).dS + dV_NN
                                src/gradient/hfscf_grad_rhf.cpp:143-167
hfscf uses Obara-Saika recursions (hfscf_osrecur.hpp); this file uses the textbook
McMurchie-Davidson scheme (Helgaker, Jorgensen, Olsen, "Molecular Electronic-Structure Theory",
9.3.1 (overlap), 9.3.4 (kinetic), 9.9.1 (Coulomb, Hermite R integrals)), vectorised over the
primitive pairs of a shell pair and over the nuclei, with the Boys function from scipy. The two
schemes give the same integrals; what fixes the numbers (component order, normalisation,
spherical order m = -l..l, half-density conventions) is hfscf's and is shared with
oracle/basis_ref.py.

Nuclear derivatives use d/dA_x G_i(x - A_x) = 2a G_{i+1} - i G_{i-1} on the bra and ket
centres and translational invariance for the operator centre: d/dC <a|1/r_C|b> =
-(d/dA + d/dB) <a|1/r_C|b>.
"""
import math

import numpy as np
from scipy.special import gamma, gammainc


def comps(l):
    """Cartesian components of angular momentum l in hfscf order (hfscf_shell.cpp:27-40)."""
    return [(l - i, i - j, j) for i in range(l + 1) for j in range(i + 1)]


def ylm(l):
    """Cartesian -> spherical matrix (ncart x 2l+1), m = -l..l, for x^l-normalised Cartesians
    (hfscf: src/math/solid_harmonics.cpp:16-109): l = 1 (y, z, x); l = 2 (sqrt3 xy, sqrt3 yz,
    zz - (xx + yy)/2, sqrt3 xz, sqrt3/2 (xx - yy))."""
    if l == 0:
        return np.ones((1, 1))
    if l == 1:
        return np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    if l == 2:
        r3 = math.sqrt(3.0)
        # Cartesian rows: xx, xy, xz, yy, yz, zz
        Y = np.zeros((6, 5))
        Y[1, 0] = r3
        Y[4, 1] = r3
        Y[5, 2], Y[0, 2], Y[3, 2] = 1.0, -0.5, -0.5
        Y[2, 3] = r3
        Y[0, 4], Y[3, 4] = 0.5 * r3, -0.5 * r3
        return Y
    raise ValueError("only s, p and d shells")


def boys(nmax, T):
    """F_n(T), n = 0..nmax, for an array T: regularised incomplete gamma, series for T < 1e-2."""
    T = np.asarray(T, dtype=float)
    out = np.empty((nmax + 1,) + T.shape)
    Ts = np.maximum(T, 1e-300)
    for n in range(nmax + 1):
        a = n + 0.5
        with np.errstate(divide="ignore", invalid="ignore", under="ignore"):
            big = gamma(a) * gammainc(a, Ts) / (2.0 * Ts ** a)
        ser = sum((-T) ** k / (math.factorial(k) * (2 * n + 2 * k + 1)) for k in range(14))
        out[n] = np.where(T < 1e-2, ser, big)
    return out


def hermite_e(imax, jmax, a, b, xab):
    """E^{ij}_t (i <= imax, j <= jmax, t <= i + j) for exponent arrays a, b and A_x - B_x:
    array [i, j, t, k] over the primitive pairs k."""
    p = a + b
    o2p = 0.5 / p
    xpa = -b / p * xab
    xpb = a / p * xab
    tmax = imax + jmax
    E = np.zeros((imax + 1, jmax + 1, tmax + 2, len(a)))
    E[0, 0, 0] = np.exp(-a * b / p * xab * xab)
    for i in range(imax + 1):
        for j in range(jmax + 1):
            if i == 0 and j == 0:
                continue
            src, x = (E[i - 1, j], xpa) if i > 0 else (E[i, j - 1], xpb)
            for t in range(i + j + 1):
                v = x * src[t] + (t + 1) * src[t + 1]
                if t > 0:
                    v = v + o2p * src[t - 1]
                E[i, j, t] = v
    return E[:, :, : tmax + 1]


def hermite_r(N, p, PC):
    """R_{tuv} (t + u + v <= N) for exponents p (shape S) and vectors PC (shape S + (3,)):
    array [t, u, v] + S, zero where t + u + v > N."""
    T = p * np.sum(PC * PC, axis=-1)
    F = boys(N, T)
    shape = T.shape
    R = {}
    for n in range(N + 1):
        R[(n, 0, 0, 0)] = (-2.0 * p) ** n * F[n]
    for s in range(1, N + 1):
        for n in range(N - s + 1):
            for t in range(s + 1):
                for u in range(s - t + 1):
                    v = s - t - u
                    if t > 0:
                        val = PC[..., 0] * R[(n + 1, t - 1, u, v)]
                        if t > 1:
                            val = val + (t - 1) * R[(n + 1, t - 2, u, v)]
                    elif u > 0:
                        val = PC[..., 1] * R[(n + 1, t, u - 1, v)]
                        if u > 1:
                            val = val + (u - 1) * R[(n + 1, t, u - 2, v)]
                    else:
                        val = PC[..., 2] * R[(n + 1, t, u, v - 1)]
                        if v > 1:
                            val = val + (v - 1) * R[(n + 1, t, u, v - 2)]
                    R[(n, t, u, v)] = val
    out = np.zeros((N + 1, N + 1, N + 1) + shape)
    for (n, t, u, v), val in R.items():
        if n == 0:
            out[t, u, v] = val
    return out


class _Pair:
    """Primitive-pair quantities of one shell pair (a on A, b on B), vectorised over pairs."""

    def __init__(self, sh, X, sa, sb, dl_a, dl_b):
        po = np.concatenate([[0], np.cumsum(sh["nprim"])])
        self.la, self.lb = int(sh["L"][sa]), int(sh["L"][sb])
        self.A = np.asarray(X[sh["atom"][sa]], float)
        self.B = np.asarray(X[sh["atom"][sb]], float)
        aa, ca = sh["alpha"][po[sa]:po[sa + 1]], sh["coef"][po[sa]:po[sa + 1]]
        ab, cb = sh["alpha"][po[sb]:po[sb + 1]], sh["coef"][po[sb]:po[sb + 1]]
        self.a = np.repeat(aa, len(ab))
        self.b = np.tile(ab, len(aa))
        self.w = np.repeat(ca, len(cb)) * np.tile(cb, len(ca))
        self.p = self.a + self.b
        self.P = (self.a[:, None] * self.A + self.b[:, None] * self.B) / self.p[:, None]
        self.E = [hermite_e(self.la + dl_a, self.lb + dl_b, self.a, self.b, self.A[d] - self.B[d]) for d in range(3)]


def _overlap_kinetic_1d(pr):
    """1-D overlap S[d][i, j, k] and kinetic T[d][i, j, k] factors, i <= la+1, j <= lb (S to lb+2)."""
    sq = np.sqrt(np.pi / pr.p)
    S = [pr.E[d][:, :, 0, :] * sq for d in range(3)]
    T = []
    for d in range(3):
        Sd = S[d]
        ni, nj = Sd.shape[0], Sd.shape[1] - 2
        Td = np.zeros((ni, nj, len(pr.p)))
        for j in range(nj):
            # -1/2 <i| d^2/dx^2 |j>, d^2 G_j = j(j-1) G_{j-2} - 2b(2j+1) G_j + 4b^2 G_{j+2}
            v = -2.0 * pr.b * (2 * j + 1) * Sd[:, j] + 4.0 * pr.b * pr.b * Sd[:, j + 2]
            if j > 1:
                v = v + j * (j - 1) * Sd[:, j - 2]
            Td[:, j] = -0.5 * v
        T.append(Td)
    return S, T


def shell_pair_blocks(sh, X, Zc, sa, sb, deriv=False):
    """Cartesian blocks of shell pair (sa, sb): S, T, V (ncart_a x ncart_b). With deriv also
    dS/dA and dT/dA (3 x na x nb; d/dB = -d/dA) and, per nucleus C, the derivatives of
    <a|-Z_C/r_C|b> with respect to A and B (3 x natom x na x nb each; d/dC = -(d/dA + d/dB))."""
    pr = _Pair(sh, X, sa, sb, 1 if deriv else 0, 3 if deriv else 2)
    la, lb = pr.la, pr.lb
    ca, cb = comps(la), comps(lb)
    S1, T1 = _overlap_kinetic_1d(pr)
    w = pr.w
    na, nb = len(ca), len(cb)
    S = np.zeros((na, nb))
    T = np.zeros((na, nb))
    for x, ea in enumerate(ca):
        for y, eb in enumerate(cb):
            s = [S1[d][ea[d], eb[d]] for d in range(3)]
            t = [T1[d][ea[d], eb[d]] for d in range(3)]
            S[x, y] = np.dot(w, s[0] * s[1] * s[2])
            T[x, y] = np.dot(w, t[0] * s[1] * s[2] + s[0] * t[1] * s[2] + s[0] * s[1] * t[2])
    # nuclear attraction: R integrals for every (primitive pair, nucleus)
    nat = len(Zc)
    N = la + lb + (1 if deriv else 0)
    PC = pr.P[:, None, :] - np.asarray(X, float)[None, :, :]
    R = hermite_r(N, np.repeat(pr.p[:, None], nat, axis=1), PC)  # [t,u,v,k,c]
    pre = -2.0 * np.pi / pr.p * w  # per primitive pair
    Zf = np.asarray(Zc, float)
    V = np.zeros((na, nb))
    for x, ea in enumerate(ca):
        for y, eb in enumerate(cb):
            V[x, y] = np.dot(np.einsum("k,kc->c", pre, _vprim(pr, R, ea, eb)), Zf)
    if not deriv:
        return S, T, V
    dS = np.zeros((3, na, nb))
    dT = np.zeros((3, na, nb))
    dVa = np.zeros((3, nat, na, nb))
    dVb = np.zeros((3, nat, na, nb))
    for x, ea in enumerate(ca):
        for y, eb in enumerate(cb):
            s = [S1[d][ea[d], eb[d]] for d in range(3)]
            t = [T1[d][ea[d], eb[d]] for d in range(3)]
            for d in range(3):
                ds = 2.0 * pr.a * S1[d][ea[d] + 1, eb[d]]
                dt = 2.0 * pr.a * T1[d][ea[d] + 1, eb[d]]
                if ea[d] > 0:
                    ds = ds - ea[d] * S1[d][ea[d] - 1, eb[d]]
                    dt = dt - ea[d] * T1[d][ea[d] - 1, eb[d]]
                e1, e2 = (d + 1) % 3, (d + 2) % 3
                dS[d, x, y] = np.dot(w, ds * s[e1] * s[e2])
                dT[d, x, y] = np.dot(w, dt * s[e1] * s[e2] + ds * (t[e1] * s[e2] + s[e1] * t[e2]))
                # V: bra shift (a exponent enters through the primitive weights, so the
                # derivative is taken per primitive pair: 2a V(a+1) - a_d V(a-1))
                va = _vshift(pr, R, pre, Zc, ea, eb, d, bra=True)
                vb = _vshift(pr, R, pre, Zc, ea, eb, d, bra=False)
                dVa[d, :, x, y] = va
                dVb[d, :, x, y] = vb
    return S, T, V, dS, dT, dVa, dVb


def _vprim(pr, R, ia, ib):
    """Primitive nuclear attraction sums sum_tuv E E E R per (primitive pair, nucleus), no prefactor."""
    ex = pr.E[0][ia[0], ib[0], : ia[0] + ib[0] + 1]
    ey = pr.E[1][ia[1], ib[1], : ia[1] + ib[1] + 1]
    ez = pr.E[2][ia[2], ib[2], : ia[2] + ib[2] + 1]
    r = R[: ex.shape[0], : ey.shape[0], : ez.shape[0]]
    return np.einsum("tk,uk,vk,tuvkc->kc", ex, ey, ez, r, optimize=True)


def _vshift(pr, R, pre, Zc, ea, eb, d, bra):
    """d/dA_d (bra) or d/dB_d (ket) of <a|-Z_C/r_C|b> for every nucleus C."""
    e, f = list(ea), list(eb)
    g = pr.a if bra else pr.b
    if bra:
        e[d] += 1
    else:
        f[d] += 1
    v = 2.0 * g[:, None] * _vprim(pr, R, tuple(e), tuple(f))
    n = ea[d] if bra else eb[d]
    if n > 0:
        e, f = list(ea), list(eb)
        if bra:
            e[d] -= 1
        else:
            f[d] -= 1
        v = v - n * _vprim(pr, R, tuple(e), tuple(f))
    return np.einsum("k,kc->c", pre, v) * np.asarray(Zc, float)


def _offsets(sh, pure):
    L = sh["L"]
    off = sh["pure_off"] if pure else sh["cart_off"]
    nf = [(2 * int(l) + 1) if pure else (int(l) + 1) * (int(l) + 2) // 2 for l in L]
    return off, nf


def _to_out(blk, la, lb, pure):
    if not pure:
        return blk
    return ylm(la).T @ blk @ ylm(lb)


def oneint(sh, pure, Z, X):
    """S, T, V (nbf x nbf) in the output basis (spherical when pure)."""
    off, nf = _offsets(sh, pure)
    n = int(max(o + k for o, k in zip(off, nf)))
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    L = sh["L"]
    ns = len(L)
    for a in range(ns):
        for b in range(a + 1):
            s, t, v = shell_pair_blocks(sh, X, Z, a, b)
            ia, ib = slice(off[a], off[a] + nf[a]), slice(off[b], off[b] + nf[b])
            for M, blk in ((S, s), (T, t), (V, v)):
                o = _to_out(blk, int(L[a]), int(L[b]), pure)
                M[ia, ib] = o
                M[ib, ia] = o.T
    return S, T, V


def enuc(Z, X):
    X = np.asarray(X, float)
    e = 0.0
    for i in range(len(Z)):
        for j in range(i):
            e += Z[i] * Z[j] / np.linalg.norm(X[i] - X[j])
    return e


def enuc_gradient(Z, X):
    X = np.asarray(X, float)
    g = np.zeros_like(X)
    for i in range(len(Z)):
        for j in range(len(Z)):
            if i != j:
                r = X[i] - X[j]
                g[i] -= Z[i] * Z[j] * r / np.linalg.norm(r) ** 3
    return g


def oneint_gradient(sh, pure, Z, X, D, W):
    """2 sum D.d(T + V) - 2 sum W.dS per atom (natom x 3), D the half density, W = D F D
    (hfscf_grad_rhf.cpp:143-167 without the two-electron and nuclear repulsion terms)."""
    off, nf = _offsets(sh, pure)
    L = sh["L"]
    atom = sh["atom"]
    nat = len(Z)
    g = np.zeros((nat, 3))
    ns = len(L)
    for a in range(ns):
        for b in range(a + 1):
            S, T, V, dS, dT, dVa, dVb = shell_pair_blocks(sh, X, Z, a, b, deriv=True)
            ia, ib = slice(off[a], off[a] + nf[a]), slice(off[b], off[b] + nf[b])
            # the (b, a) block is the transpose of (a, b): count off-diagonal shell pairs twice
            fac = 1.0 if a == b else 2.0
            Dab = D[ia, ib]
            Wab = W[ia, ib]
            A, B = int(atom[a]), int(atom[b])
            for d in range(3):
                ds = _to_out(dS[d], int(L[a]), int(L[b]), pure)
                dt = _to_out(dT[d], int(L[a]), int(L[b]), pure)
                # S and T depend on A - B only: d/dB = -d/dA
                c = fac * (2.0 * np.sum(Dab * dt) - 2.0 * np.sum(Wab * ds))
                g[A, d] += c
                g[B, d] -= c
                for C in range(nat):
                    va = _to_out(dVa[d, C], int(L[a]), int(L[b]), pure)
                    vb = _to_out(dVb[d, C], int(L[a]), int(L[b]), pure)
                    ca = fac * 2.0 * np.sum(Dab * va)
                    cb = fac * 2.0 * np.sum(Dab * vb)
                    g[A, d] += ca
                    g[B, d] += cb
                    g[C, d] -= ca + cb  # operator centre: translational invariance
    return g


def rhf_gradient(sh, pure, Z, X, D, F, g2e):
    """Total RHF gradient: the two-electron part g2e (sum [2 D_pq D_rs - D_pr D_qs] (pq|rs)'),
    the one-electron terms with W = D F D, and the nuclear repulsion gradient."""
    W = D @ F @ D
    return g2e + oneint_gradient(sh, pure, Z, X, D, W) + enuc_gradient(Z, X)
