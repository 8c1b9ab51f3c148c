"""GPU parity of the packed ERI vector (hfg_eri_vector) against the oracle's stored ERIs, which
restate hfscf's stored repulsion vector (canonical index_ijkl order, hfscf's screening).

Bar: 1e-12 absolute per element (the ERI bar), every element including the screened zeros."""
import numpy as np
import pytest

import hfscf_amd as H
import oracle
from conftest import golden
from oracle import basis_ref

pytestmark = pytest.mark.gpu


def _check(mol, screen):
    ctx = H.GpuContext(mol, schwarz=screen)
    v = ctx.eri_vector()
    o = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=screen)
    o.build_eri()
    r = o.eri_vector()
    assert v.shape == r.shape
    err = float(np.abs(v - r).max())
    print(f"nbf {mol.nbf} screen {screen}: {len(v)} ERIs, max |d| {err:.1e}, "
          f"{int((v != 0).sum())} / {int((r != 0).sum())} nonzero", flush=True)
    assert err < 1e-12  # includes the elements one side screened to zero (values below 5e-15)
    # a second call on the same context gives the same vector
    assert np.array_equal(ctx.eri_vector(), v)
    return v


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
@pytest.mark.parametrize("screen", [True, False])
def test_eri_vector_water(basis, screen):
    _check(H.Molecule.from_xyz(golden("h2o.xyz"), basis), screen)


def test_eri_vector_water_dimer_and_digestion():
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    mol = H.Molecule.from_arrays(Z[:6], X[:6], "cc-pvdz")
    v = _check(mol, True)
    # J from the exported vector (hfscf's stored-ERI digestion) against the direct J/K build
    n = mol.nbf
    rng = np.random.default_rng(3)
    A = rng.standard_normal((n, n))
    D = 0.1 * (A + A.T)
    i, j = np.tril_indices(n)
    ij = i * (i + 1) // 2 + j
    npair = n * (n + 1) // 2
    P = np.zeros((npair, npair))
    a, b = np.tril_indices(npair)
    P[a, b] = v[a * (a + 1) // 2 + b]
    P = P + P.T - np.diag(np.diag(P))
    w = np.where(i == j, 1.0, 2.0) * D[i, j]
    Jv = np.zeros((n, n))
    Jv[i, j] = P[ij] @ w
    Jv = Jv + Jv.T - np.diag(np.diag(Jv))
    J, K = H.GpuContext(mol).build_jk(D)
    print(f"J from the vector vs direct build: {np.abs(Jv - J).max():.1e}", flush=True)
    assert np.abs(Jv - J).max() < 1e-11
