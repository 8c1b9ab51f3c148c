"""GPU parity of the C-ABI entry points that the J/K and gradient tests do not reach directly:
the device Boys functions (hfg_boys_probe) across their branch points, the Schwarz diagonal
(hfg_schwarz) and the two-electron energy (hfg_energy, the forward call whose backward is the
adjoint gradient).

References: scipy's regularised incomplete gamma (oracle/md_eri.py) and the closed forms
F_m(0) = 1/(2m+1), F_0(T) = sqrt(pi/T) erf(sqrt T)/2; the oracle's screen matrix and its
stored-ERI J/K. Bars: Boys 5e-14 relative; Schwarz diagonal 1e-13 relative with the same
exactly-zero entries; energy 1e-11 Eh.
"""
import math

import numpy as np
import pytest
from scipy.special import erf

import hfscf_amd as H
import oracle
from conftest import golden
from oracle import basis_ref, md_eri

pytestmark = pytest.mark.gpu

# grid points of the table (k/32), the Taylor midpoints, the closed-form switch at 0.05, the
# table end at 36 and the asymptotic branch
T_POINTS = [0.0, 1e-14, 1e-8, 1e-3, 1.0 / 64, 0.049, 0.05, 0.051, 0.5, 1.0, 2.0 + 1.0 / 64, 5.0, 17.3, 35.98, 35.99,
            36.0, 36.02, 50.0, 120.0, 1000.0]


def _ref(m, T):
    if T == 0.0:
        return 1.0 / (2 * m + 1)
    return md_eri.boys(m, T)


@pytest.fixture(scope="module")
def water_ctx():
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), "sto-3g")
    return H.GpuContext(mol)


@pytest.mark.parametrize("m", [0, 1, 2, 4, 8, 12, 17])
def test_boys_table_branches(water_ctx, m):
    F = water_ctx.boys(m, T_POINTS)
    worst = 0.0
    for i, T in enumerate(T_POINTS):
        for k in range(m + 1):
            ref = _ref(k, T)
            worst = max(worst, abs(F[i, k] - ref) / ref)
    print(f"Boys table m <= {m}: max relative error {worst:.2e}", flush=True)
    assert worst < 5e-14


def test_boys_closed_forms(water_ctx):
    F = water_ctx.boys(1, T_POINTS, closed_form=True)
    for i, T in enumerate(T_POINTS):
        f0 = 1.0 if T == 0.0 else 0.5 * math.sqrt(math.pi / T) * erf(math.sqrt(T))
        assert abs(F[i, 0] - f0) <= 5e-15 * f0, (T, F[i, 0], f0)
        assert abs(F[i, 1] - _ref(1, T)) <= 2e-14 * _ref(1, T), (T, F[i, 1])


def test_boys_rejects_bad_arguments(water_ctx):
    with pytest.raises(RuntimeError):
        water_ctx.boys(18, [1.0])
    with pytest.raises(RuntimeError):
        water_ctx.boys(2, [1.0], closed_form=True)
    with pytest.raises(RuntimeError):
        water_ctx.boys(0, [-1.0])


@pytest.mark.parametrize("basis,dimer", [("6-31g*", False), ("cc-pvdz", True)])
def test_schwarz_matches_oracle_screen(basis, dimer):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    if dimer:
        Z = np.concatenate([Z, Z])
        X = np.vstack([X, X + np.array([9.0, 1.0, -2.0])])
    mol = H.Molecule.from_arrays(Z, X, basis)
    S = H.GpuContext(mol).schwarz()
    ref = oracle.Oracle(mol.shells(), X, mol.pure, 1e-14, screen=True).screen()
    assert S.shape == ref.shape
    assert np.all((S == 0) == (ref == 0))
    nz = ref != 0
    rel = np.abs(S[nz] - ref[nz]) / np.abs(ref[nz])
    print(f"{basis} {'dimer' if dimer else 'water'}: {int((~nz).sum())} zero entries, max rel {rel.max():.2e}",
          flush=True)
    assert rel.max() < 1e-13


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_energy_matches_oracle(basis):
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    mol = H.Molecule.from_arrays(Z[:6], X[:6], basis)
    rng = np.random.default_rng(11)
    A = rng.standard_normal((mol.nbf, mol.nbf))
    D = 0.05 * (A + A.T)
    e = H.GpuContext(mol).energy(D)
    J, K = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=True).jk(D)
    ref = float(np.sum(D * (2 * J - K)))
    print(f"(H2O)2 {basis}: E_2e {e:.12f}, oracle {ref:.12f}, diff {e - ref:.2e}", flush=True)
    assert abs(e - ref) < 1e-11
