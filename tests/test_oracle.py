"""Pins of the oracle (CPU, no GPU): every check compares it with something independent.

* Boys F_m(T): scipy's regularised incomplete gamma, closed forms F_m(0) = 1/(2m+1),
  F_0(T) = sqrt(pi/T) erf(sqrt T)/2, and the large-T asymptote.
* ERIs: an independent McMurchie-Davidson evaluator (oracle/md_eri.py), the closed-form
  (ss|ss), 8-fold permutational symmetry and the Schwarz inequality.
* Derivative ERIs: central finite differences of the oracle's own ERIs under a centre shift,
  and translational invariance d/dA + d/dB + d/dC + d/dD = 0.
* SCF: the product's host SCF loop and a numpy restatement of hfscf's loop reach the same
  energy, and water/STO-3G at the geometry of the Crawford programming project reproduces
  its published E_nuc and (to 1e-6, basis-data provenance) total energy.
"""
import itertools
import math

import numpy as np
import pytest
from scipy.special import erf

import hfscf_amd as H
import oracle
from conftest import golden
from oracle import basis_ref, md_eri, scf_ref


@pytest.mark.parametrize("m", [0, 1, 2, 5, 9, 13])
def test_boys_against_scipy(m):
    for T in np.concatenate([[1e-7, 1e-3, 0.1, 0.5], np.linspace(1, 60, 25)]):
        ref = md_eri.boys(m, float(T))
        assert abs(oracle.boys(m, float(T)) - ref) <= 2e-14 * max(1.0, abs(ref)), (m, T)


def test_boys_closed_forms():
    for T in [0.01, 0.3, 2.0, 7.5, 30.0]:
        assert abs(oracle.boys(0, T) - 0.5 * math.sqrt(math.pi / T) * erf(math.sqrt(T))) < 1e-15
    # hfscf floors T at 1e-8, so F_m(0) comes out as F_m(1e-8)
    for m in range(8):
        assert abs(oracle.boys(m, 0.0) - (1.0 / (2 * m + 1) - 1e-8 / (2 * m + 3))) < 1e-14
    T = 80.0
    for m in range(6):
        asym = math.prod(range(2 * m - 1, 0, -2)) / 2 ** (m + 1) * math.sqrt(math.pi / T ** (2 * m + 1))
        assert abs(oracle.boys(m, T) - asym) < 1e-14 * asym


def _water(basis, screen=False):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    sh, pure, nbf = basis_ref.shells_for(Z, basis)
    return sh, X, pure, oracle.Oracle(sh, X, pure, 1e-14, screen=screen)


def _md_block(sh, X, pure, q):
    ref = md_eri.contracted_eri(sh, X, *q)
    if not pure:
        return ref
    r3 = math.sqrt(3.0)
    Y = {0: np.ones((1, 1)), 1: np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], float), 2: np.zeros((6, 5))}
    Y[2][1, 0] = r3
    Y[2][4, 1] = r3
    Y[2][5, 2] = 1.0
    Y[2][0, 2] = Y[2][3, 2] = -0.5
    Y[2][2, 3] = r3
    Y[2][0, 4] = r3 / 2
    Y[2][3, 4] = -r3 / 2
    Ls = [int(sh["L"][s]) for s in q]
    return np.einsum("abcd,ai,bj,ck,dl->ijkl", ref, Y[Ls[0]], Y[Ls[1]], Y[Ls[2]], Y[Ls[3]])


def test_eri_sto3g_all_unique_against_md():
    sh, X, pure, o = _water("sto-3g")
    ns = len(sh["L"])
    worst = 0.0
    for i, j, k, l in itertools.product(range(ns), repeat=4):
        if j > i or l > k or (i * (i + 1) // 2 + j) < (k * (k + 1) // 2 + l):
            continue
        d = np.abs(o.shell_quartet(i, j, k, l) - _md_block(sh, X, pure, (i, j, k, l))).max()
        worst = max(worst, d)
    assert worst < 1e-13


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_eri_d_shells_against_md(basis):
    sh, X, pure, o = _water(basis)
    L = sh["L"]
    d = int(np.where(L == 2)[0][0])
    p = int(np.where(L == 1)[0][0])
    h = len(L) - 1
    for q in [(d, d, d, d), (d, p, d, 0), (p, d, h, d), (d, 0, p, p), (h, d, h, p)]:
        assert np.abs(o.shell_quartet(*q) - _md_block(sh, X, pure, q)).max() < 1e-13, q


def test_ss_closed_form():
    # two s primitives on different centres: (ss|ss) = 2 pi^2.5/(pq sqrt(p+q)) K_ab K_cd F0(T)
    sh = dict(L=np.array([0, 0], np.int32), atom=np.array([0, 1], np.int32), nprim=np.array([1, 1], np.int32),
              alpha=np.array([0.8, 1.3]), coef_raw=np.array([1.0, 1.0]))
    X = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 1.1]])
    o = oracle.Oracle(sh, X, False, 1e-14, screen=False)
    a, b = 0.8, 1.3
    na, nb = (2 * a / math.pi) ** 0.75, (2 * b / math.pi) ** 0.75
    p = q = a + b
    rab2 = float(np.sum(X[1] ** 2))
    Kab = math.exp(-a * b / p * rab2)
    val = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * Kab * Kab * (na * nb) ** 2
    assert abs(o.shell_quartet(0, 1, 0, 1)[0, 0, 0, 0] - val) < 1e-14


def test_eri_symmetry_and_schwarz():
    sh, X, pure, o = _water("cc-pvdz")
    rng = np.random.default_rng(3)
    ns = len(sh["L"])
    for _ in range(25):
        i, j, k, l = (int(x) for x in rng.integers(0, ns, 4))
        b = o.shell_quartet(i, j, k, l)
        assert np.allclose(b, o.shell_quartet(j, i, k, l).transpose(1, 0, 2, 3), atol=1e-14, rtol=0)
        assert np.allclose(b, o.shell_quartet(k, l, i, j).transpose(2, 3, 0, 1), atol=1e-14, rtol=0)
        assert np.allclose(b, o.shell_quartet(i, j, l, k).transpose(0, 1, 3, 2), atol=1e-14, rtol=0)
        d1 = np.einsum("abab->ab", o.shell_quartet(i, j, i, j))
        d2 = np.einsum("abab->ab", o.shell_quartet(k, l, k, l))
        bound = np.sqrt(np.abs(d1[:, :, None, None] * d2[None, None, :, :]))
        assert np.all(np.abs(b) <= bound + 1e-13)


def test_deriv1_finite_difference_and_translation():
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    sh, pure, _ = basis_ref.shells_for(Z, "6-31g*")  # Cartesian, s p d
    L = sh["L"]
    quartets = [(5, 2, 4, 0), (5, 3, 7, 1), (2, 3, 5, 8)]  # four distinct shells each
    h = 1e-5
    for q in quartets:
        o = oracle.Oracle(sh, X, False, 1e-14, screen=False)
        der = o.shell_quartet_deriv1(*q)
        total = der[0:3] + der[3:6] + der[6:9]
        # translational invariance determines d/dD; check it against a finite difference
        for centre in range(4):
            atom = int(sh["atom"][q[centre]])
            # move only this shell's centre: use a one-off atom by duplicating the geometry
            for d in range(3):
                Xp, Xm = X.copy(), X.copy()
                sh2 = {k: v.copy() for k, v in sh.items()}
                new_atom = len(X)
                sh2["atom"][q[centre]] = new_atom
                Xp = np.vstack([X, X[atom]])
                Xm = np.vstack([X, X[atom]])
                Xp[new_atom, d] += h
                Xm[new_atom, d] -= h
                op = oracle.Oracle(sh2, Xp, False, 1e-14, screen=False)
                om = oracle.Oracle(sh2, Xm, False, 1e-14, screen=False)
                fd = (op.shell_quartet(*q) - om.shell_quartet(*q)) / (2 * h)
                an = der[3 * centre + d] if centre < 3 else -total[d]
                assert np.abs(fd - an).max() < 2e-7, (q, centre, d, np.abs(fd - an).max())


def test_host_scf_matches_numpy_restatement():
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), "cc-pvdz")
    o = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=True)
    S, T, V = mol.oneint()
    r = H.rhf(mol, jk=o.jk, diis_range=6)
    ref = scf_ref.rhf(S, T + V, mol.nelec // 2, o.jk, diis_range=6)
    assert r["converged"] and ref["converged"]
    assert abs(r["energy"] - (ref["energy_elec"] + mol.enuc())) < 1e-9


def test_crawford_water_sto3g():
    # Crawford group programming project #3 geometry (bohr); E_nuc = 8.002367061810450,
    # E_SCF = -74.942079928192 (published; basis data provenance bounds this to ~1e-6)
    Xb = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                   [-1.638036840407, 1.136548822547, 0.0]])
    mol = H.Molecule.from_arrays(np.array([8, 1, 1]), Xb, "sto-3g")
    assert abs(mol.enuc() - 8.002367061810450) < 1e-12
    o = oracle.Oracle(mol.shells(), Xb, mol.pure, 1e-14, screen=True)
    r = H.rhf(mol, jk=o.jk, diis_range=6)
    assert abs(r["energy"] - (-74.942079928192)) < 1e-6


def test_cfg1_fixture_reproduces():
    f = np.load(golden("cfg1_h2o_sto3g.npz"))
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), "sto-3g")
    o = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=False)
    r = H.rhf(mol, jk=o.jk, diis_range=6, energy_tol=1e-10, rms_tol=1e-10)
    assert abs(r["energy"] - float(f["energy"])) < 1e-11


def test_eri_pair_matrix_positive_semidefinite():
    # (ij|kl) as a matrix over function pairs is a Gram matrix of the pair densities under the
    # Coulomb metric, so it has no negative eigenvalue; a wrong sign or index in the recursion
    # breaks this (cc-pVDZ water, 24 functions, every quartet from the packed ERI vector)
    sh, X, pure, o = _water("cc-pvdz")
    o.build_eri()
    v = o.eri_vector()
    n = o.nbf
    M = np.zeros((n * n, n * n))

    def idx2(a, b):
        return a * (a + 1) // 2 + b if a >= b else b * (b + 1) // 2 + a

    for i in range(n):
        for j in range(n):
            ij = idx2(i, j)
            for k in range(n):
                for l in range(n):
                    M[i * n + j, k * n + l] = v[idx2(ij, idx2(k, l))]
    w = np.linalg.eigvalsh(M)
    assert w.min() > -1e-12 * w.max(), (w.min(), w.max())
