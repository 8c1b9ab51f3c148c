"""Write the oracle's golden vectors used by the GPU parity tests (run on a CPU host).

    python tests/golden/make_fixtures.py [cfg1 cfg2 cfg3 grad h2o16]

Each fixture is a small .npz of inputs and expected outputs produced by the oracle
(the CPU restatement of hfscf's two-electron algorithm) through the host RHF loop:
  cfg1_h2o_sto3g.npz     energies per iteration, first and last Fock matrices, density
  cfg2_h2o8_ccpvdz.npz   converged SCF energies (loop and final build); J/K rows for a seeded
                         random density
  cfg3_c12h26_631gss.npz the same for C12H26 / 6-31G**
  grad_h2o.npz           two-electron and total RHF gradients of water (3 bases)
  grad_h2o16.npz         (H2O)16 / cc-pVDZ converged energy and analytic RHF gradient
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import hfscf_amd as H  # noqa: E402
import oracle  # noqa: E402

ROWS = [0, 1, 7, 33, 64, 101, 150, 191]


def seeded_density(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return 0.05 * (A + A.T)


def cfg1():
    mol = H.Molecule.from_xyz(os.path.join(HERE, "h2o.xyz"), "sto-3g")
    orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=False)
    S, T, V = mol.oneint()
    Hc = T + V
    focks = []

    def jk(D):
        J, K = orc.jk(D)
        focks.append(Hc + 2 * J - K)
        return J, K

    r = H.rhf(mol, jk=jk, diis_range=6, energy_tol=1e-10, rms_tol=1e-10)
    np.savez_compressed(os.path.join(HERE, "cfg1_h2o_sto3g.npz"), energy=r["energy"], energies=r["energies"],
                        F_first=focks[0], F_last=focks[-1], D=r["D"], H=Hc, S=S)
    print("cfg1", r["energy"], r["iterations"])


def big(name, xyz, basis, rows_seed=21, diis=8):
    mol = H.Molecule.from_xyz(os.path.join(HERE, xyz), basis)
    orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=True)
    t = time.time()
    orc.build_eri()
    print(name, "eri build", time.time() - t, flush=True)
    # converged past the 1e-10 of the workload definition so that energy_final (one more full
    # build at the converged density) is fixed to well below the 1e-9 Eh parity bar
    r = H.rhf(mol, jk=orc.jk, diis_range=diis, energy_tol=1e-12, rms_tol=1e-11)
    print(name, "energy", r["energy"], r["energy_final"], r["iterations"], flush=True)
    D = seeded_density(mol.nbf, rows_seed)
    J, K = orc.jk(D)
    rows = [k for k in ROWS if k < mol.nbf]
    np.savez_compressed(os.path.join(HERE, name), energy=r["energy"], energy_final=r["energy_final"],
                        iterations=r["iterations"], seed=rows_seed, rows=np.array(rows), J_rows=J[rows],
                        K_rows=K[rows], EJ=np.sum(D * J), EK=np.sum(D * K))


def grad():
    out = {}
    for basis in ["sto-3g", "cc-pvdz", "6-31g*"]:
        mol = H.Molecule.from_xyz(os.path.join(HERE, "h2o.xyz"), basis)
        orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=False)
        r = H.rhf(mol, jk=orc.jk, diis_range=6, energy_tol=1e-12, rms_tol=1e-11)
        g2 = orc.grad2e(r["D"])
        g = H.rhf_gradient(mol, r["D"], r["F"], jk=orc.jk, grad2e=orc.grad2e)
        key = basis.replace("*", "s").replace("-", "")
        out[f"{key}_D"] = r["D"]
        out[f"{key}_F"] = r["F"]
        out[f"{key}_g2e"] = g2
        out[f"{key}_g"] = g
        out[f"{key}_E"] = r["energy"]
        print("grad", basis, r["energy"], np.abs(g.sum(0)).max())
    np.savez_compressed(os.path.join(HERE, "grad_h2o.npz"), **out)


def h2o16():
    # (H2O)16 / cc-pVDZ: converged SCF (stored-ERI oracle, about 22 GB of ERIs) and the oracle's
    # analytic RHF gradient at that density (on-the-fly derivative ERIs)
    mol = H.Molecule.from_xyz(os.path.join(HERE, "h2o_16.xyz"), "cc-pvdz")
    orc = oracle.Oracle(mol.shells(), mol.geometry()[1], mol.pure, 1e-14, screen=True)
    t = time.time()
    orc.build_eri()
    print("h2o16 eri build", time.time() - t, flush=True)
    r = H.rhf(mol, jk=orc.jk, diis_range=8, energy_tol=1e-12, rms_tol=1e-10)
    print("h2o16 energy", r["energy"], r["energy_final"], r["iterations"], flush=True)
    # the energies first (the analytic gradient below takes hours of CPU)
    np.savez_compressed(os.path.join(HERE, "energy_h2o16.npz"), energy=r["energy"], energy_final=r["energy_final"],
                        iterations=r["iterations"])
    t = time.time()
    g = H.rhf_gradient(mol, r["D"], r["F"], jk=orc.jk, grad2e=orc.grad2e)
    print("h2o16 gradient", time.time() - t, np.abs(g.sum(0)).max(), flush=True)
    np.savez_compressed(os.path.join(HERE, "grad_h2o16.npz"), energy=r["energy"], energy_final=r["energy_final"],
                        iterations=r["iterations"], g=g)


if __name__ == "__main__":
    which = sys.argv[1:] or ["cfg1", "cfg2", "grad"]
    if "cfg1" in which:
        cfg1()
    if "grad" in which:
        grad()
    if "cfg2" in which:
        big("cfg2_h2o8_ccpvdz.npz", "h2o_8.xyz", "cc-pvdz")
    if "cfg3" in which:
        big("cfg3_c12h26_631gss.npz", "c12h26.xyz", "6-31g**")
    if "h2o16" in which:
        h2o16()
