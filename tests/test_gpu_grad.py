"""GPU parity of the adjoint two-electron gradient kernel and of the full RHF gradient.

Pins: the oracle's on-the-fly derivative-ERI contraction (committed vectors for water and a
live run for the water dimer), central finite differences of the GPU SCF energy, and the
translational / rotational invariance of the total gradient. Tolerances: two-electron part
1e-10 Eh/bohr against the oracle, total gradient 1e-8 Eh/bohr against the oracle-assembled
gradient, 2e-6 Eh/bohr against finite differences (step 1e-4 bohr, O(h^2) truncation).
"""
import numpy as np
import pytest

import hfscf_amd as H
import oracle
from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_grad2e_water_fixture(basis):
    f = np.load(golden("grad_h2o.npz"))
    key = basis.replace("*", "s").replace("-", "")
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), basis)
    ctx = H.GpuContext(mol)
    g = ctx.energy_backward(f[f"{key}_D"])
    d = np.abs(g - f[f"{key}_g2e"]).max()
    print(f"{basis}: max |g2e - oracle| = {d:.2e}", flush=True)
    assert d < 1e-10
    gt = H.rhf_gradient(mol, f[f"{key}_D"], f[f"{key}_F"])
    dt = np.abs(gt - f[f"{key}_g"]).max()
    print(f"{basis}: max |g - oracle total| = {dt:.2e}", flush=True)
    assert dt < 1e-8


def test_grad2e_dimer_live_oracle():
    from oracle import basis_ref
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    mol = H.Molecule.from_arrays(Z[:6], X[:6], "cc-pvdz")
    rng = np.random.default_rng(9)
    A = rng.standard_normal((mol.nbf, mol.nbf))
    D = 0.05 * (A + A.T)
    g = H.GpuContext(mol).energy_backward(D)
    orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=False)
    gr = orc.grad2e(D)
    d = np.abs(g - gr).max()
    print(f"(H2O)2 cc-pVDZ random D: max |g2e - oracle| = {d:.2e}", flush=True)
    assert d < 1e-9


def _energy(mol, xyz, **opts):
    mol.set_geometry(xyz)
    r = H.rhf(mol, diis_range=8, energy_tol=1e-12, rms_tol=1e-11, **opts)
    assert r["converged"]
    return r["energy_final"]


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_gradient_finite_difference_water(basis):
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), basis)
    X0 = mol.geometry()[1].copy()
    r = H.rhf(mol, diis_range=8, energy_tol=1e-12, rms_tol=1e-11)
    g = H.rhf_gradient(mol, r["D"], r["F"])
    assert np.abs(g.sum(0)).max() < 1e-9
    torque = np.cross(X0, g).sum(0)
    assert np.abs(torque).max() < 1e-8
    h = 1e-4
    for atom, d in [(0, 2), (1, 0), (2, 2)]:
        Xp, Xm = X0.copy(), X0.copy()
        Xp[atom, d] += h
        Xm[atom, d] -= h
        fd = (_energy(mol, Xp) - _energy(mol, Xm)) / (2 * h)
        print(f"{basis} atom {atom} dir {d}: analytic {g[atom, d]:.10f} fd {fd:.10f} diff {g[atom, d] - fd:.2e}",
              flush=True)
        assert abs(g[atom, d] - fd) < 2e-6
    mol.set_geometry(X0)


def test_h2o16_gradient_fd_mixed_and_oracle():
    # differentiation workload: (H2O)16 / cc-pVDZ (384 functions, 48 atoms). FP64 analytic
    # gradient against central differences of the GPU energy on 6 seeded coordinates (seed 4,
    # h = 1e-4 bohr), the mixed-precision gradient against FP64 (bar 1e-5 Eh/bohr), the oracle's
    # SCF energy (bar 1e-9 Eh) and the oracle's analytic gradient when its fixture is present
    # (bar 1e-8 Eh/bohr)
    mol = H.Molecule.from_xyz(golden("h2o_16.xyz"), "cc-pvdz")
    assert mol.nbf == 384 and mol.natom == 48
    X0 = mol.geometry()[1].copy()
    opts = dict(diis_range=8, energy_tol=1e-12, rms_tol=1e-10)
    r = H.rhf(mol, **opts)
    assert r["converged"]
    g = H.rhf_gradient(mol, r["D"], r["F"])
    assert np.abs(g.sum(0)).max() < 1e-8
    rm = H.rhf(mol, precision=H.MIXED, incremental=1, **opts)
    gm = H.rhf_gradient(mol, rm["D"], rm["F"], precision=H.MIXED)
    gk = H.rhf_gradient(mol, r["D"], r["F"], precision=H.MIXED)  # mixed kernels at the FP64 density
    print(f"(H2O)16: E {r['energy_final']:.10f}, mixed dE {rm['energy_final'] - r['energy_final']:.2e}, "
          f"mixed gradient max |dg| {np.abs(gk - g).max():.2e} at the FP64 density, "
          f"{np.abs(gm - g).max():.2e} after the mixed SCF", flush=True)
    assert abs(rm["energy_final"] - r["energy_final"]) <= 1e-6
    # the precision-mode bar on the gradient kernels (same density); after the mixed SCF the
    # density itself differs (the FP32 Fock error, about 6e-6, shifts the self-consistent solution),
    # which adds 1e-5 to 4e-5 (measured)
    assert np.abs(gk - g).max() <= 1e-5
    assert np.abs(gm - g).max() <= 1e-4
    import os
    fe = np.load(golden("energy_h2o16.npz"))  # the oracle's SCF (stored ERIs) on this cluster
    print(f"(H2O)16 vs oracle: dE {r['energy_final'] - float(fe['energy_final']):.2e} "
          f"({int(fe['iterations'])} oracle iterations)", flush=True)
    assert abs(r["energy_final"] - float(fe["energy_final"])) < 1e-9
    path = golden("grad_h2o16.npz")
    if os.path.exists(path):
        f = np.load(path)
        print(f"(H2O)16 vs oracle: dE {r['energy_final'] - float(f['energy_final']):.2e}, "
              f"max |dg| {np.abs(g - f['g']).max():.2e}", flush=True)
        assert abs(r["energy_final"] - float(f["energy_final"])) < 1e-9
        assert np.abs(g - f["g"]).max() < 1e-8
    rng = np.random.default_rng(4)
    h = 1e-4
    for c in sorted(rng.choice(mol.natom * 3, size=6, replace=False).tolist()):
        atom, d = divmod(c, 3)
        Xp, Xm = X0.copy(), X0.copy()
        Xp[atom, d] += h
        Xm[atom, d] -= h
        fd = (_energy(mol, Xp) - _energy(mol, Xm)) / (2 * h)
        print(f"(H2O)16 atom {atom} dir {d}: analytic {g[atom, d]:.10f} fd {fd:.10f} diff {g[atom, d] - fd:.2e}",
              flush=True)
        assert abs(g[atom, d] - fd) < 2e-6
    mol.set_geometry(X0)
