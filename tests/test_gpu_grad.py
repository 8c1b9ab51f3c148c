"""GPU parity of the adjoint two-electron gradient kernel and of the full RHF gradient.

Pins: the oracle's on-the-fly derivative-ERI contraction (committed vectors for water and a
live run for the water dimer), central finite differences of the GPU SCF energy, and the
translational / rotational invariance of the total gradient. Tolerances: two-electron part
1e-10 Eh/bohr against the oracle, total gradient 1e-8 Eh/bohr against the oracle's (water and
the (H2O)16 differentiation workload), central differences (step 1e-4 bohr) of unscreened
energies 1e-8 Eh/bohr for water and 1e-7 Eh/bohr for (H2O)16, mixed precision 1e-5 Eh/bohr.
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


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_grad2e_cooperative_kernels_from_l4(basis):
    # the gradient classes from L = 4 on the cooperative gradient kernel, against the oracle
    f = np.load(golden("grad_h2o.npz"))
    key = basis.replace("*", "s").replace("-", "")
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), basis)
    g = H.GpuContext(mol, coop_min_l=4).energy_backward(f[f"{key}_D"])
    d = np.abs(g - f[f"{key}_g2e"]).max()
    print(f"{basis}, coop from L=4: max |g2e - oracle| = {d:.2e}", flush=True)
    assert d < 1e-10


def test_grad2e_dimer_live_oracle():
    from oracle import basis_ref
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    mol = H.Molecule.from_arrays(Z[:6], X[:6], "cc-pvdz")
    rng = np.random.default_rng(9)
    A = rng.standard_normal((mol.nbf, mol.nbf))
    D = 0.05 * (A + A.T)
    g = H.GpuContext(mol).energy_backward(D)
    gc = H.GpuContext(mol, coop_min_l=4).energy_backward(D)
    orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=False)
    gr = orc.grad2e(D)
    d, dc = np.abs(g - gr).max(), np.abs(gc - gr).max()
    print(f"(H2O)2 cc-pVDZ random D: max |g2e - oracle| = {d:.2e} (cooperative from L=4: {dc:.2e})", flush=True)
    assert d < 1e-9 and dc < 1e-9


def _energy(mol, xyz, **opts):
    mol.set_geometry(xyz)
    r = H.rhf(mol, diis_range=8, energy_tol=1e-12, rms_tol=1e-11, **opts)
    assert r["converged"]
    return r["energy_final"]


def _fd(mol, X0, atom, d, h=1e-4):
    # unscreened energies at the displaced points: the Schwarz screen's shell and element skips
    # change with the geometry and would put steps into the energy surface
    Xp, Xm = X0.copy(), X0.copy()
    Xp[atom, d] += h
    Xm[atom, d] -= h
    try:
        return (_energy(mol, Xp, schwarz=0) - _energy(mol, Xm, schwarz=0)) / (2 * h)
    finally:
        mol.set_geometry(X0)


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_gradient_finite_difference_water(basis):
    mol = H.Molecule.from_xyz(golden("h2o.xyz"), basis)
    X0 = mol.geometry()[1].copy()
    r = H.rhf(mol, diis_range=8, energy_tol=1e-12, rms_tol=1e-11)
    g = H.rhf_gradient(mol, r["D"], r["F"])
    assert np.abs(g.sum(0)).max() < 1e-9
    torque = np.cross(X0, g).sum(0)
    assert np.abs(torque).max() < 1e-8
    for atom, d in [(0, 2), (1, 0), (2, 2)]:
        fd = _fd(mol, X0, atom, d)
        print(f"{basis} atom {atom} dir {d}: analytic {g[atom, d]:.10f} fd {fd:.10f} diff {g[atom, d] - fd:.2e}",
              flush=True)
        assert abs(g[atom, d] - fd) < 1e-8


# differentiation workload: (H2O)16 / cc-pVDZ (384 functions, 48 atoms)
H2O16_OPTS = dict(diis_range=8, energy_tol=1e-12, rms_tol=1e-11)


@pytest.fixture(scope="module")
def h2o16():
    mol = H.Molecule.from_xyz(golden("h2o_16.xyz"), "cc-pvdz")
    assert mol.nbf == 384 and mol.natom == 48
    r = H.rhf(mol, **H2O16_OPTS)
    assert r["converged"]
    g = H.rhf_gradient(mol, r["D"], r["F"])
    return mol, r, g


def test_h2o16_gradient_against_oracle(h2o16):
    # the oracle alone (its one-electron integrals, stored ERIs, SCF loop and on-the-fly
    # derivative-ERI contraction) on the same geometry; bars 1e-9 Eh and 1e-8 Eh/bohr
    mol, r, g = h2o16
    f = np.load(golden("grad_h2o16.npz"))
    de = r["energy_final"] - float(f["energy_final"])
    dg = np.abs(g - f["g"]).max()
    print(f"(H2O)16 vs oracle: E {r['energy_final']:.10f} dE {de:.2e}, max |dg| {dg:.2e}, "
          f"sum g {np.abs(g.sum(0)).max():.1e}", flush=True)
    assert abs(de) < 1e-9
    assert dg < 1e-8
    assert np.abs(g.sum(0)).max() < 1e-8
    # rotational invariance: no torque (the oracle's own gradient has 4.9e-8 at its convergence)
    torque = np.abs(np.cross(mol.geometry()[1], g).sum(0)).max()
    print(f"(H2O)16 torque {torque:.1e}", flush=True)
    assert torque < 2e-7
    fe = np.load(golden("energy_h2o16.npz"))
    assert abs(r["energy_final"] - float(fe["energy_final"])) < 1e-9


def test_h2o16_mixed_gradient(h2o16):
    # mixed precision: FP32 density-difference builds, an FP64 finish (from rms 1e-6) and the FP32
    # adjoint kernels; bars |dE| 1e-6 Eh (reached: about 1e-10) and 1e-5 Eh/bohr end to end
    mol, r, g = h2o16
    rm = H.rhf(mol, precision=H.MIXED, incremental=1, **H2O16_OPTS)
    assert rm["converged"]
    gm = H.rhf_gradient(mol, rm["D"], rm["F"], precision=H.MIXED)
    gk = H.rhf_gradient(mol, r["D"], r["F"], precision=H.MIXED)  # mixed kernels at the FP64 density
    de = rm["energy_final"] - r["energy_final"]
    print(f"(H2O)16 mixed: dE {de:.2e}, max |dg| {np.abs(gk - g).max():.2e} at the FP64 density, "
          f"{np.abs(gm - g).max():.2e} end to end ({rm['iterations']} iterations)", flush=True)
    assert abs(de) <= 1e-6
    assert np.abs(gk - g).max() <= 1e-5
    assert np.abs(gm - g).max() <= 1e-5


def _rotation(seed):
    q = np.random.default_rng(seed).standard_normal(4)
    a, b, c, d = q / np.linalg.norm(q)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


@pytest.mark.parametrize("basis", ["cc-pvdz", "6-31g*"])
def test_energy_and_gradient_rotation_translation_invariance(basis):
    # the water dimer rotated by a seeded rotation and shifted: the same energy, the gradient
    # rotated with it (unscreened builds: the screen's element skips follow Cartesian axes)
    from oracle import basis_ref
    Z, X = basis_ref.read_xyz(golden("h2o_8.xyz"))
    Z, X = Z[:6], X[:6]
    R, t = _rotation(11), np.array([0.7, -1.3, 2.1])
    res = []
    for xyz in (X, X @ R.T + t):
        mol = H.Molecule.from_arrays(Z, xyz, basis)
        r = H.rhf(mol, diis_range=8, energy_tol=1e-12, rms_tol=1e-11, schwarz=0)
        assert r["converged"]
        res.append((r["energy_final"], H.rhf_gradient(mol, r["D"], r["F"])))
    (e0, g0), (e1, g1) = res
    dg = np.abs(g1 - g0 @ R.T).max()
    print(f"{basis}: dE {e1 - e0:.2e}, |g(R X + t) - R g(X)| {dg:.2e}", flush=True)
    assert abs(e1 - e0) < 1e-10
    assert dg < 1e-8


def _h2o16_coords():
    rng = np.random.default_rng(4)
    return [divmod(c, 3) for c in sorted(rng.choice(48 * 3, size=6, replace=False).tolist())]


@pytest.mark.parametrize("part", [0, 1, 2])
def test_h2o16_gradient_finite_difference(h2o16, part):
    # central differences (h = 1e-4 bohr) of unscreened GPU energies on 6 seeded coordinates
    # (seed 4), two per test; bar 1e-7 Eh/bohr
    mol, r, g = h2o16
    X0 = mol.geometry()[1].copy()
    for atom, d in _h2o16_coords()[2 * part:2 * part + 2]:
        fd = _fd(mol, X0, atom, d)
        print(f"(H2O)16 atom {atom} dir {d}: analytic {g[atom, d]:.10f} fd {fd:.10f} diff {g[atom, d] - fd:.2e}",
              flush=True)
        assert abs(g[atom, d] - fd) < 1e-7
