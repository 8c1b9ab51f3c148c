"""GPU parity of the one-electron integrals (S, T, V) and their gradient terms against the
oracle's numpy restatement (McMurchie-Davidson; oracle/oneint_ref.py) and the host layer.

Tolerances: matrix elements 1e-12 absolute, gradient terms 1e-11 Eh/bohr (the host layer's bars
against the same oracle)."""
import numpy as np
import pytest

import hfscf_amd as H
from conftest import golden
from oracle import basis_ref
from oracle import oneint_ref as O

pytestmark = pytest.mark.gpu


def _water(basis):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    X = X + np.array([[0.1, -0.05, 0.02], [0.0, 0.03, -0.04], [0.02, 0.0, 0.05]])
    sh, pure, n = basis_ref.shells_for(Z, basis, None)
    return Z, X, sh, pure, n


def _sym(n, seed):
    A = np.random.default_rng(seed).standard_normal((n, n))
    return 0.1 * (A + A.T)


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_oneint_matches_oracle(basis):
    Z, X, sh, pure, n = _water(basis)
    mol = H.Molecule.from_arrays(Z, X, basis)
    S, T, V = H.GpuContext(mol).oneint()
    So, To, Vo = O.oneint(sh, pure, Z, X)
    err = [float(np.abs(a - b).max()) for a, b in ((S, So), (T, To), (V, Vo))]
    print(f"{basis}: |dS| {err[0]:.1e} |dT| {err[1]:.1e} |dV| {err[2]:.1e}", flush=True)
    assert max(err) < 1e-12


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_oneint_gradient_matches_oracle(basis):
    Z, X, sh, pure, n = _water(basis)
    mol = H.Molecule.from_arrays(Z, X, basis)
    D, Q = _sym(n, 5), _sym(n, 6)
    g = H.GpuContext(mol).oneint_gradient(D, Q)
    go = O.oneint_gradient(sh, pure, Z, X, D, Q)
    print(f"{basis}: |dg| {np.abs(g - go).max():.1e}", flush=True)
    assert np.abs(g - go).max() < 1e-11


def test_oneint_h2o8_matches_host():
    # 24 atoms, 96 shells: every shell-pair class of cc-pVDZ and many nuclei per lane
    mol = H.Molecule.from_xyz(golden("h2o_8.xyz"), "cc-pvdz")
    ctx = H.GpuContext(mol)
    S, T, V = ctx.oneint()
    Sh, Th, Vh = mol.oneint()
    err = max(float(np.abs(a - b).max()) for a, b in ((S, Sh), (T, Th), (V, Vh)))
    D, Q = _sym(mol.nbf, 7), _sym(mol.nbf, 8)
    Z, X = mol.geometry()
    g = ctx.oneint_gradient(D, Q) + O.enuc_gradient(Z, X)
    gh = mol.oneint_gradient(D, Q)
    print(f"(H2O)8: |dSTV| {err:.1e}, |dg| {np.abs(g - gh).max():.1e}", flush=True)
    assert err < 1e-12
    assert np.abs(g - gh).max() < 1e-10
    # a context without nuclear charges refuses
    with pytest.raises(RuntimeError):
        H.GpuContext([mol]).oneint()
