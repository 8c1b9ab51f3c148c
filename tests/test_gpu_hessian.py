"""GPU parity of the two-electron Hessian at a fixed density (hfg_hessian_2e) against the oracle's
restatement (second-derivative ERIs by the shift rule applied twice, pinned by finite differences
of its gradient in tests/test_oracle_hessian.py).

Tolerance: 1e-8 Eh/bohr^2 absolute on every element of the Coulomb and exchange parts (the
gradient bar, 1e-8 Eh/bohr, per bohr). At (H2O)4 / cc-pVDZ, size-independent properties:
symmetry and translational invariance (each row sums to zero over the atoms per direction)."""
import numpy as np
import pytest

import hfscf_amd as H
import oracle
from conftest import golden
from oracle import basis_ref

pytestmark = pytest.mark.gpu


def _density(n, nocc, seed):
    C = np.random.default_rng(seed).standard_normal((n, nocc)) / np.sqrt(n)
    return C @ C.T


def _compare(Z, X, basis, seed, label):
    mol = H.Molecule.from_arrays(Z, X, basis)
    D = _density(mol.nbf, 5, seed)
    hc, hx = H.GpuContext(mol).hessian_2e(D)
    rc, rx = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=False).hess2e(D)
    ec, ex = float(np.abs(hc - rc).max()), float(np.abs(hx - rx).max())
    print(f"{label}: |dHC| {ec:.2e} |dHX| {ex:.2e} (max |HC| {np.abs(rc).max():.2e}, |HX| {np.abs(rx).max():.2e})",
          flush=True)
    assert ec < 1e-8 and ex < 1e-8


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g*", "cc-pvdz"])
def test_hessian_2e_water_matches_oracle(basis):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    X = X + np.array([[0.1, -0.05, 0.02], [0.0, 0.03, -0.04], [0.02, 0.0, 0.05]])
    _compare(Z, X, basis, 3, f"water {basis}")


def test_hessian_2e_water_dimer_matches_oracle():
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    _compare(Z[:6], X[:6], "cc-pvdz", 5, "(H2O)2 cc-pvdz")


def test_hessian_2e_h2o4_invariants():
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    mol = H.Molecule.from_arrays(Z[:12], X[:12], "cc-pvdz")
    D = _density(mol.nbf, 20, 9)
    hc, hx = H.GpuContext(mol).hessian_2e(D)
    n3 = hc.shape[0]
    for name, M in (("HC", hc), ("HX", hx)):
        sym = float(np.abs(M - M.T).max())
        trans = float(np.abs(M.reshape(n3, -1, 3).sum(axis=1)).max())
        print(f"(H2O)4 {name}: max {np.abs(M).max():.2e} asym {sym:.1e} translation {trans:.1e}", flush=True)
        assert sym < 1e-9 and trans < 1e-8



def test_hessian_2e_rejects_bad_arguments():
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    mol = H.Molecule.from_arrays(Z, X, "sto-3g")
    ctx = H.GpuContext(mol)
    with pytest.raises(ValueError):
        ctx.hessian_2e(np.zeros((mol.nbf + 1, mol.nbf + 1)))
    n3 = 3 * len(Z)
    # null outputs are refused by the entry point itself, with the context's message
    rc = H.hfgpu().hfg_hessian_2e(ctx._h, H._ptr(np.zeros((mol.nbf, mol.nbf))), None, None)
    assert rc != 0
    hc, hx = ctx.hessian_2e(np.zeros((mol.nbf, mol.nbf)))  # a zero density gives zero Hessians
    assert hc.shape == (n3, n3) and not hc.any() and not hx.any()
