"""Pins of the oracle's screening semantics (CPU only).

hfscf screens in three places (hfscf: src/integrals/hfscf_oseri.cpp:36-46 shell skip on the
mean of sqrt|S_pq S_rs| over the quartet, :152-158 element skip and the Cartesian diagonal taken
from the screen matrix; src/integrals/hfscf_oseri_screen.cpp:78 the screen matrix's own
primitive cutoff |pfac| < 1e-16). These tests check the oracle's screening against what does not
depend on reading those lines the same way twice:
* the screen matrix against an independent McMurchie-Davidson evaluation of (pq|pq) with the
  same primitive cutoff rule: identical to 1e-14 relative, including every entry that the rule
  makes exactly zero (and that is nonzero without it);
* at a coarse integral_tol the screened J and K differ from the unscreened ones, and (there and
  at the production tolerance) by no more than the Cauchy-Schwarz bound of the dropped
  integrals, sum |D_rs| sqrt((pq|pq)(rs|rs)) with the exact diagonal, plus the change of the
  Cartesian diagonal, which must equal the screen matrix.
"""
import numpy as np
import pytest

import oracle
from conftest import golden
from oracle import basis_ref, md_eri


def _dimer(basis, pure=False, shift=(9.0, 1.0, -2.0)):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    Z2 = np.concatenate([Z, Z])
    X2 = np.vstack([X, X + np.array(shift)])
    sh, pure, n = basis_ref.shells_for(Z2, basis, pure)
    return Z2, X2, sh, pure, n


def _unpack(v, n):
    """Full (ij|kl) from hfscf's packed index_ijkl vector."""
    i, j, k, l = np.meshgrid(*(np.arange(n),) * 4, indexing="ij")
    ij = np.where(i > j, i * (i + 1) // 2 + j, j * (j + 1) // 2 + i)
    kl = np.where(k > l, k * (k + 1) // 2 + l, l * (l + 1) // 2 + k)
    idx = np.where(ij > kl, ij * (ij + 1) // 2 + kl, kl * (kl + 1) // 2 + ij)
    return v[idx]


def test_screen_matrix_primitive_cut_against_md():
    Z, X, sh, pure, n = _dimer("6-31g*")
    o = oracle.Oracle(sh, X, False, 1e-14, screen=True)
    scr = o.screen()
    ns = len(sh["L"])
    co = sh["cart_off"]
    zeros_cut = zeros_oracle = 0
    worst = 0.0
    for a in range(ns // 2):
        for b in range(ns // 2, ns):  # shells of different monomers: the cut removes whole entries
            d = np.einsum("abab->ab", md_eri.contracted_eri(sh, X, a, b, a, b, prim_cut=1e-16))
            d0 = np.einsum("abab->ab", md_eri.contracted_eri(sh, X, a, b, a, b))
            s = scr[co[a]:co[a] + d.shape[0], co[b]:co[b] + d.shape[1]]
            zeros_cut += int(np.sum((d == 0) & (d0 != 0)))
            zeros_oracle += int(np.sum(s == 0))
            assert np.all((s == 0) == (d == 0))
            if d.any():
                worst = max(worst, float(np.abs(s - d).max() / np.abs(d).max()))
    assert zeros_cut > 0 and zeros_oracle == zeros_cut
    assert worst < 1e-13


def _jk(eri4, D):
    J = np.einsum("pqrs,rs->pq", eri4, D)
    K = np.einsum("prqs,rs->pq", eri4, D)
    return J, K


@pytest.mark.parametrize("tol", [1e-8, 1e-14])
def test_screening_error_within_schwarz_bound(tol):
    # a coarse integral_tol (screen_tol = 1e-6) so that screening bites, and the production one
    Z, X, sh, pure, n = _dimer("6-31g*", shift=(5.5, 0.5, -1.0))
    o_s = oracle.Oracle(sh, X, False, tol, screen=True)
    o_u = oracle.Oracle(sh, X, False, tol, screen=False)
    o_s.build_eri()
    o_u.build_eri()
    I_s, I_u = _unpack(o_s.eri_vector(), n), _unpack(o_u.eri_vector(), n)
    scr = o_s.screen()
    rng = np.random.default_rng(2)
    A = rng.standard_normal((n, n))
    D = 0.1 * (A + A.T)
    Js, Ks = o_s.jk(D)
    Ju, Ku = o_u.jk(D)
    # the oracle's digestion agrees with a plain contraction of its stored integrals
    for (J, K), I in (((Js, Ks), I_s), ((Ju, Ku), I_u)):
        Jr, Kr = _jk(I, D)
        assert np.abs(J - Jr).max() < 1e-12 and np.abs(K - Kr).max() < 1e-12
    dropped = (I_s == 0) & (I_u != 0)
    if tol > 1e-10:
        assert dropped.sum() > 0
    # the only other change: the Cartesian diagonal (pq|pq) is the screen matrix's value
    changed = (I_s != I_u) & ~dropped
    p, q, r, t = np.nonzero(changed)
    assert np.all(((p == r) & (q == t)) | ((p == t) & (q == r)))  # (pq|pq) under 8-fold symmetry
    assert np.abs(I_s[p, q, p, q] - scr[p, q]).max(initial=0.0) < 1e-15
    # Cauchy-Schwarz with the exact diagonal (no primitive cutoff at all): the screen matrix's
    # own 1e-16 cut zeroes (pq|pq) of distant pairs, so its bound would be 0 for integrals that
    # the screen drops although they are not (up to about 1e-9 here)
    o_x = oracle.Oracle(sh, X, False, 1e-30, screen=False)
    o_x.build_eri()
    I_x = _unpack(o_x.eri_vector(), n)
    sq = np.sqrt(np.abs(np.einsum("pqpq->pq", I_x)))
    slack = 100.0 * tol / 1e4  # the unscreened integrals still drop primitives below tol/1e4
    bound4 = np.where(dropped, sq[:, :, None, None] * sq[None, None, :, :] + slack, 0.0)
    assert np.all(np.abs(I_u[dropped]) <= bound4[dropped])
    bound4 = bound4 + np.where(changed, np.abs(I_s - I_u), 0.0)
    bJ = np.einsum("pqrs,rs->pq", bound4, np.abs(D)) + 1e-13
    bK = np.einsum("prqs,rs->pq", bound4, np.abs(D)) + 1e-13
    dJ, dK = np.abs(Js - Ju), np.abs(Ks - Ku)
    if tol > 1e-10:
        assert dJ.max() > 1e-12 and dK.max() > 1e-12
    assert np.all(dJ <= bJ) and np.all(dK <= bK)
