"""Pins of the oracle's one-electron integrals (oracle/oneint_ref.py), CPU only, and the
product's host one-electron layer checked against them.

Each oracle check compares with something independent of the McMurchie-Davidson code:
* closed forms for two s primitives (overlap, kinetic, nuclear attraction);
* unit overlap diagonal (spherical functions; s, p and the xx/yy/zz Cartesian d components),
  1/3 for the xy/xz/yz components (hfscf normalises the x^L component only);
* kinetic energy as -1/2 the Laplacian of the overlap with respect to the ket centre (finite
  differences, Richardson-extrapolated);
* nuclear attraction as the limit of the oracle's Obara-Saika ERIs (ref_oracle.cpp, pinned
  against an independent McMurchie-Davidson ERI code in test_oracle.py) against a unit-charge
  s Gaussian of exponent 1e14 on each nucleus: the error of that limit is pi (ab)(C) / zeta;
* nuclear derivatives by central finite differences of the one-electron energy terms;
* water/STO-3G at the Crawford programming project geometry, from the oracle alone (one-electron
  integrals here, ERIs and the SCF loop of the oracle): the published total energy.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

import hfscf_amd as H
import oracle
from conftest import golden
from oracle import basis_ref, scf_ref
from oracle import oneint_ref as O


def _water(basis, pure=None, jitter=False):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    if jitter:
        X = X + np.array([[0.1, -0.05, 0.02], [0.0, 0.03, -0.04], [0.02, 0.0, 0.05]])
    sh, pure, n = basis_ref.shells_for(Z, basis, pure)
    return Z, X, sh, pure, n


def test_ss_closed_forms():
    a, b = 0.8, 1.3
    A, B, C = np.array([0.0, 0.0, 0.0]), np.array([0.3, -0.4, 1.1]), np.array([-0.7, 0.2, 0.5])
    sh = dict(L=np.array([0, 0]), atom=np.array([0, 1]), nprim=np.array([1, 1]), alpha=np.array([a, b]),
              coef=np.array([(2 * a / math.pi) ** 0.75, (2 * b / math.pi) ** 0.75]))
    X = np.array([A, B, C])
    Zc = np.array([0.0, 0.0, 3.0])
    S, T, V = O.shell_pair_blocks(sh, X, Zc, 1, 0)
    p, mu = a + b, a * b / (a + b)
    R2 = float(np.sum((A - B) ** 2))
    s_ref = sh["coef"][0] * sh["coef"][1] * (math.pi / p) ** 1.5 * math.exp(-mu * R2)
    t_ref = mu * (3.0 - 2.0 * mu * R2) * s_ref
    P = (a * A + b * B) / p
    Tb = p * float(np.sum((P - C) ** 2))
    f0 = 0.5 * math.sqrt(math.pi / Tb) * erf(math.sqrt(Tb))
    v_ref = -3.0 * 2.0 * math.pi / p * sh["coef"][0] * sh["coef"][1] * math.exp(-mu * R2) * f0
    assert abs(S[0, 0] - s_ref) < 1e-15
    assert abs(T[0, 0] - t_ref) < 1e-14
    assert abs(V[0, 0] - v_ref) < 1e-14


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_overlap_diagonal_normalisation(basis):
    Z, X, sh, pure, n = _water(basis)
    S, _, _ = O.oneint(sh, pure, Z, X)
    want = np.ones(n)
    if not pure:
        # x^L-only normalisation: xy, xz, yz of a Cartesian d shell have norm 1/3
        for s, l in enumerate(sh["L"]):
            if l == 2:
                o = int(sh["cart_off"][s])
                want[[o + 1, o + 2, o + 4]] = 1.0 / 3.0
    assert np.abs(np.diag(S) - want).max() < 2e-15
    assert np.linalg.eigvalsh(S).min() > 0


def test_kinetic_is_laplacian_of_overlap():
    Z, X, sh, _, _ = _water("6-31g*", pure=False, jitter=True)
    ns = len(sh["L"])

    def lap(a, b, h):
        sh2 = {k: v.copy() for k, v in sh.items()}
        sh2["atom"][b] = len(Z)  # the ket shell alone on its own centre
        Z2 = np.append(Z, 0)
        out = 0.0
        for d in range(3):
            v = []
            for s in (-1, 0, 1):
                X2 = np.vstack([X, X[sh["atom"][b]]])
                X2[-1, d] += s * h
                v.append(O.shell_pair_blocks(sh2, X2, Z2, a, b)[0])
            out = out + (v[0] - 2 * v[1] + v[2]) / h ** 2
        return out

    worst = 0.0
    for a in range(ns):
        for b in range(ns):
            if a == b:
                continue
            T = O.shell_pair_blocks(sh, X, Z, a, b)[1]
            h = 1e-3
            L2 = (4.0 * lap(a, b, h / 2) - lap(a, b, h)) / 3.0  # Richardson: O(h^4)
            worst = max(worst, np.abs(T + 0.5 * L2).max() / max(1.0, np.abs(T).max()))
    assert worst < 2e-8, worst


def test_nuclear_attraction_is_point_charge_limit_of_eris():
    Z, X, sh, _, _ = _water("6-31g*", pure=False, jitter=True)
    ns, npr, nat = len(sh["L"]), len(sh["alpha"]), len(Z)
    zeta = 1e14
    sh2 = {k: np.concatenate([sh[k], np.zeros(nat, dtype=sh[k].dtype)]) for k in ("L", "nprim", "atom", "alpha",
                                                                              "coef_raw")}
    sh2["nprim"][ns:] = 1
    sh2["atom"][ns:] = np.arange(nat)
    sh2["alpha"][npr:] = zeta / 2  # product of two: a unit-charge Gaussian of exponent zeta
    sh2["coef_raw"][npr:] = 1.0
    o = oracle.Oracle(sh2, X, False, 1e-30, screen=False)
    worst = 0.0
    for a in range(ns):
        for b in range(a + 1):
            V = O.shell_pair_blocks(sh, X, Z, a, b)[2]
            ref = sum(-Z[C] * o.shell_quartet(a, b, ns + C, ns + C)[:, :, 0, 0] for C in range(nat))
            worst = max(worst, np.abs(V - ref).max())
    assert worst < 1e-10, worst


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_oneint_gradient_finite_difference(basis):
    Z, X, sh, pure, n = _water(basis, jitter=True)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n))
    D = 0.1 * (A + A.T)
    A = rng.standard_normal((n, n))
    W = 0.1 * (A + A.T)

    def f(Xc):
        S, T, V = O.oneint(sh, pure, Z, Xc)
        return 2 * np.sum(D * (T + V)) - 2 * np.sum(W * S) + O.enuc(Z, Xc)

    g = O.oneint_gradient(sh, pure, Z, X, D, W) + O.enuc_gradient(Z, X)
    h = 1e-4
    for a in range(3):
        for d in range(3):
            Xp, Xm = X.copy(), X.copy()
            Xp[a, d] += h
            Xm[a, d] -= h
            fd = (f(Xp) - f(Xm)) / (2 * h)
            assert abs(fd - g[a, d]) < 1e-7, (a, d, fd, g[a, d])
    assert np.abs(g.sum(0)).max() < 1e-13


def test_crawford_water_sto3g_oracle_only():
    # Crawford group programming project #3 (bohr): E_nuc = 8.002367061810450,
    # E_SCF = -74.942079928192 (published; basis-data provenance bounds this to ~1e-6)
    Xb = np.array([[0.0, -0.143225816552, 0.0], [1.638036840407, 1.136548822547, 0.0],
                   [-1.638036840407, 1.136548822547, 0.0]])
    Z = np.array([8, 1, 1])
    sh, pure, n = basis_ref.shells_for(Z, "sto-3g")
    S, T, V = O.oneint(sh, pure, Z, Xb)
    assert abs(O.enuc(Z, Xb) - 8.002367061810450) < 1e-12
    o = oracle.Oracle(sh, Xb, pure, 1e-14, screen=True)
    r = scf_ref.rhf(S, T + V, 5, o.jk, diis_range=6)
    assert r["converged"]
    assert abs(r["energy_elec"] + O.enuc(Z, Xb) - (-74.942079928192)) < 1e-6


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_product_oneint_matches_oracle(basis):
    Z, X, sh, pure, n = _water(basis, jitter=True)
    mol = H.Molecule.from_arrays(Z, X, basis)
    S, T, V = mol.oneint()
    So, To, Vo = O.oneint(sh, pure, Z, X)
    for a, b in ((S, So), (T, To), (V, Vo)):
        assert np.abs(a - b).max() < 1e-12
    assert abs(mol.enuc() - O.enuc(Z, X)) < 1e-12


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_product_oneint_gradient_matches_oracle(basis):
    Z, X, sh, pure, n = _water(basis, jitter=True)
    mol = H.Molecule.from_arrays(Z, X, basis)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((n, n))
    D = 0.1 * (A + A.T)
    A = rng.standard_normal((n, n))
    Q = 0.1 * (A + A.T)
    g = mol.oneint_gradient(D, Q)
    go = O.oneint_gradient(sh, pure, Z, X, D, Q) + O.enuc_gradient(Z, X)
    assert np.abs(g - go).max() < 1e-11
