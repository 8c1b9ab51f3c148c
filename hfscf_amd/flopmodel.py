"""Algorithmic FLOP model v1 of the RHF two-electron hot path (executable definition).

Per unique contracted shell quartet (la lb|lc ld) with K primitive quartets and L = la+lb+lc+ld:

    F = K * [30 (setup) + (30 + 3L) (Boys seed + downward) + VRR + 2 N_con] + HRR + SPH + 12 na nb nc nd

VRR   2 FLOP per Obara-Saika term building (e0|f0)^(m) for |e| <= la+lb, |f| <= lc+ld,
      m <= L - |e| - |f|; bra targets are built first with 2 + 2[e_d - 1 > 0] terms, ket
      targets with 2 + 2[f_d - 1 > 0] + [e_d > 0] terms; d is the target's first nonzero
      Cartesian direction.
N_con sum_{A=la}^{la+lb} ncart(A) * sum_{C=lc}^{lc+ld} ncart(C)
HRR   2 FLOP per bra-HRR and ket-HRR target.
SPH   2 nnz(l) * (product of the other three current dimensions) per spherical index with
      l >= 2 (nnz(d) = 8), applied to the indices in order a, b, c, d.
Digestion: 6 FMA (12 FLOP) per function quartet.

Gradient model v1: VRR at (la+lb+1, lc+ld+1, L+1), N_con over the shifted ranges, 3x HRR and
SPH, and 26 FLOP per function quartet.

The C++ twin used at run time is hfscf_amd/csrc/gpu/flopmodel.hpp.
"""


def ncart(l):
    return (l + 1) * (l + 2) // 2


def comps(l):
    return [(l - i, i - j, j) for i in range(l + 1) for j in range(i + 1)]


def first_dir(e):
    return 0 if e[0] > 0 else (1 if e[1] > 0 else 2)


def vrr_flops(lab, lcd, lt):
    f = 0
    for le in range(lab + 1):
        for e in comps(le):
            if le >= 1:
                d = first_dir(e)
                f += 2 * (2 + 2 * (e[d] - 1 > 0)) * (lt - le + 1)
            for lf in range(1, lcd + 1):
                mcount = lt - le - lf + 1
                if mcount <= 0:
                    continue
                for g in comps(lf):
                    d = first_dir(g)
                    f += 2 * (2 + 2 * (g[d] - 1 > 0) + (e[d] > 0)) * mcount
    return f


def quartet_flops(la, lb, lc, ld, K, pure, gradient=False):
    sh = 1 if gradient else 0
    L = la + lb + lc + ld + sh
    lab, lcd = la + lb + sh, lc + ld + sh
    ncon = sum(ncart(a) for a in range(la, la + lb + sh + 1)) * sum(ncart(c) for c in range(lc, lc + ld + sh + 1))
    per_prim = 30 + (30 + 3 * L) + vrr_flops(lab, lcd, L) + 2 * ncon
    nf_ket = sum(ncart(c) for c in range(lc, lc + ld + 1))
    bra_t = sum(sum(ncart(a) for a in range(la, la + lb - kb + 1)) * ncart(kb) for kb in range(1, lb + 1))
    ket_t = sum(sum(ncart(c) for c in range(lc, lc + ld - kd + 1)) * ncart(kd) for kd in range(1, ld + 1))
    hrr = 2 * (bra_t * nf_ket + ket_t * ncart(la) * ncart(lb))
    dims = [ncart(la), ncart(lb), ncart(lc), ncart(ld)]
    ls = [la, lb, lc, ld]
    sph = 0
    if pure:
        for t in range(4):
            if ls[t] >= 2:
                other = 1
                for u in range(4):
                    if u != t:
                        other *= dims[u]
                sph += 2 * 8 * other
            if ls[t] >= 1:
                dims[t] = 2 * ls[t] + 1
    nfq = 1
    for l in ls:
        nfq *= (2 * l + 1 if l >= 1 else 1) if pure else ncart(l)
    if gradient:
        return K * per_prim + 3 * (hrr + sph) + 26 * nfq
    return K * per_prim + hrr + sph + 12 * nfq


if __name__ == "__main__":
    print("VRR (pp|pp)", vrr_flops(2, 2, 4), "(dd|dd)", vrr_flops(4, 4, 8))
