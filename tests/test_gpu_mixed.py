"""Mixed-precision mode and density-weighted screening of incremental builds, against FP64.

Mixed mode evaluates and digests shell quartets in FP32 (Boys values, pair data, Schwarz
bounds, LDS tiles, slabs and the gather in FP64). Its tolerances follow the targets of the
mode: |dE| <= 1e-6 Eh for the SCF energy and <= 1e-5 Eh/bohr for the gradient; the tests also
check that the FP32 path really ran (results differ from FP64 beyond FP64 rounding).
"""
import os

import numpy as np
import pytest

import hfscf_amd as H

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _mol(xyz, basis):
    return H.Molecule.from_xyz(os.path.join(GOLD, xyz), basis)


def _rand_density(n, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return scale * (A + A.T)


def test_mixed_jk_close_to_fp64():
    mol = _mol("h2o_8.xyz", "cc-pvdz")
    D = _rand_density(mol.nbf, 11)
    J, K = H.GpuContext(mol).build_jk(D)
    Jm, Km = H.GpuContext(mol, precision=H.MIXED).build_jk(D)
    dj, dk = np.abs(Jm - J).max(), np.abs(Km - K).max()
    print(f"mixed vs FP64: |dJ| {dj:.2e} (max |J| {np.abs(J).max():.2e}), |dK| {dk:.2e}", flush=True)
    assert dj < 1e-5 * np.abs(J).max() and dk < 1e-5 * np.abs(K).max()
    assert max(dj, dk) > 1e-11  # the FP32 kernels ran


def test_set_precision_switches_and_rejects_unknown():
    mol = _mol("h2o.xyz", "cc-pvdz")
    ctx = H.GpuContext(mol)
    D = _rand_density(mol.nbf, 3)
    J64, _ = ctx.build_jk(D)
    ctx.set_precision(H.MIXED)
    Jm, _ = ctx.build_jk(D)
    ctx.set_precision(H.FP64)
    J64b, _ = ctx.build_jk(D)
    assert np.abs(J64b - J64).max() < 1e-13
    assert 1e-12 < np.abs(Jm - J64).max() < 1e-5
    with pytest.raises(RuntimeError):
        ctx.set_precision(7)


@pytest.mark.parametrize("precision", [H.FP64, H.MIXED])
def test_incremental_density_screen(precision):
    # incremental build on a small density change with the density-weighted screen on,
    # against a full FP64 build of the same density
    mol = _mol("h2o_8.xyz", "cc-pvdz")
    S, T, V = mol.oneint()
    Hc = T + V
    D1 = _rand_density(mol.nbf, 5)
    D2 = D1 + 1e-4 * _rand_density(mol.nbf, 6)
    ctx = H.GpuContext(mol, density_tol=1e-11)
    ctx.fock(Hc, D1)
    ctx.set_precision(precision)
    Finc = ctx.fock(Hc, D2, incremental=True)
    Ffull = H.GpuContext(mol).fock(Hc, D2)
    err = np.abs(Finc - Ffull).max()
    print(f"incremental (precision {precision}) vs full: {err:.2e}", flush=True)
    assert err < (1e-9 if precision == H.FP64 else 1e-8)


@pytest.mark.parametrize("xyz,basis", [("h2o.xyz", "cc-pvdz"), ("h2o_8.xyz", "cc-pvdz")])
def test_mixed_scf_energy(xyz, basis):
    mol = _mol(xyz, basis)
    ref = H.rhf(mol, diis_range=8)
    mix = H.rhf(mol, diis_range=8, precision=H.MIXED, incremental=1)
    de = mix["energy_final"] - ref["energy_final"]
    print(f"{xyz}/{basis}: mixed SCF dE {de:.2e} ({mix['iterations']} vs {ref['iterations']} iterations, "
          f"Fock {mix['fock_seconds']:.3f} s vs {ref['fock_seconds']:.3f} s)", flush=True)
    assert mix["converged"] and ref["converged"]
    assert abs(de) <= 1e-6
    if xyz == "h2o_8.xyz":
        f = np.load(os.path.join(GOLD, "cfg2_h2o8_ccpvdz.npz"))
        key = "energy_final" if "energy_final" in f else "energy"
        assert abs(mix["energy_final"] - float(f[key])) <= 1e-6


def test_mixed_scf_c12h26():
    # precision-mode workload: C12H26 / 6-31G** (Cartesian d, 310 functions)
    mol = _mol("c12h26.xyz", "6-31g**")
    ref = H.rhf(mol, diis_range=8)
    # FP32 density-difference builds leave an error of order 1e-8 in the accumulated Fock matrix,
    # which bounds how far the DIIS error can fall: the mixed run stops at rms 1e-7 and its
    # energy is that of the final FP64 build
    mix = H.rhf(mol, diis_range=8, precision=H.MIXED, incremental=1, energy_tol=1e-9, rms_tol=1e-7)
    de = mix["energy_final"] - ref["energy_final"]
    print(f"C12H26: E {ref['energy_final']:.10f} ({ref['iterations']} iterations, converged {ref['converged']}), "
          f"mixed dE {de:.2e} ({mix['iterations']} iterations, converged {mix['converged']})", flush=True)
    assert mix["converged"] and ref["converged"]
    assert abs(de) <= 1e-6
    path = os.path.join(GOLD, "cfg3_c12h26_631gss.npz")
    if os.path.exists(path):
        f = np.load(path)
        key = "energy_final" if "energy_final" in f else "energy"
        print(f"C12H26 vs oracle: FP64 {ref['energy_final'] - float(f[key]):.2e}, "
              f"mixed {mix['energy_final'] - float(f[key]):.2e}", flush=True)
        assert abs(ref["energy_final"] - float(f[key])) <= 1e-9
        assert abs(mix["energy_final"] - float(f[key])) <= 1e-6


@pytest.mark.parametrize("basis", ["6-31g*", "cc-pvdz"])
def test_mixed_gradient(basis):
    mol = _mol("h2o.xyz", basis)
    scf = H.rhf(mol, diis_range=6, energy_tol=1e-12, rms_tol=1e-11)
    g64 = H.GpuContext(mol).energy_backward(scf["D"])
    gm = H.GpuContext(mol, precision=H.MIXED).energy_backward(scf["D"])
    err = np.abs(gm - g64).max()
    print(f"water/{basis}: mixed two-electron gradient error {err:.2e}", flush=True)
    assert err <= 1e-5
    assert err > 1e-13  # the FP32 kernels ran
    assert np.abs(gm.sum(0)).max() < 1e-8  # translational invariance holds per quartet in FP64 sums


def test_mixed_gradient_h2o8():
    mol = _mol("h2o_8.xyz", "cc-pvdz")
    scf = H.rhf(mol, diis_range=8)
    g64 = H.GpuContext(mol).energy_backward(scf["D"])
    gm = H.GpuContext(mol, precision=H.MIXED).energy_backward(scf["D"])
    err = np.abs(gm - g64).max()
    print(f"(H2O)8: mixed two-electron gradient error {err:.2e}", flush=True)
    assert err <= 1e-5
