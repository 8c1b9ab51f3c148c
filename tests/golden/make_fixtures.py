"""Write the oracle's golden vectors used by the GPU parity tests (run on a CPU host).

    python tests/golden/make_fixtures.py [cfg1 cfg2 cfg3 grad h2o16]

Each fixture is a small .npz of inputs and expected outputs produced by the oracle
(the CPU restatement of hfscf's two-electron algorithm) through the host RHF loop:
  cfg1_h2o_sto3g.npz     energies per iteration, first and last Fock matrices, density
  cfg2_h2o8_ccpvdz.npz   converged SCF energies (loop and final build); J/K rows for a seeded
                         random density
  cfg3_c12h26_631gss.npz the same for C12H26 / 6-31G**
  grad_h2o.npz           two-electron and total RHF gradients of water (3 bases)
  energy_h2o16.npz       (H2O)16 / cc-pVDZ converged energy, from the oracle alone
  grad_h2o16.npz         the same and the analytic RHF gradient (total and two-electron part)
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
    """Water / STO-3G, conventional SCF with stored ERIs (no screening), DIIS 6, 1e-10: from the
    oracle alone. energies are electronic energies per iteration (hfscf_hf_solver.cpp:244-252);
    F_first / F_last are the Fock matrices of the first and last builds and D_first / D_last
    the densities they were built from."""
    from oracle import basis_ref, oneint_ref, scf_ref

    Z, X = basis_ref.read_xyz(os.path.join(HERE, "h2o.xyz"))
    sh, pure, n = basis_ref.shells_for(Z, "sto-3g")
    S, T, V = oneint_ref.oneint(sh, pure, Z, X)
    Hc = T + V
    orc = oracle.Oracle(sh, X, pure, 1e-14, screen=False)
    dens, focks = [], []

    def jk(D):
        J, K = orc.jk(D)
        dens.append(D.copy())
        focks.append(Hc + 2 * J - K)
        return J, K

    r = scf_ref.rhf(S, Hc, int(sum(Z)) // 2, jk, diis_range=6, energy_tol=1e-10, rms_tol=1e-10)
    energy = r["energy_elec"] + oneint_ref.enuc(Z, X)
    np.savez_compressed(os.path.join(HERE, "cfg1_h2o_sto3g.npz"), energy=energy, energies=np.array(r["energies"]),
                        F_first=focks[0], F_last=focks[-1], D_first=dens[0], D_last=dens[-1], D=r["D"], H=Hc, S=S)
    print("cfg1", energy, r["iterations"])


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
    """(H2O)16 / cc-pVDZ from the oracle alone (no product code): basis and ordering
    (oracle/basis_ref.py), one-electron integrals and gradient terms (oracle/oneint_ref.py),
    stored ERIs and the two-electron gradient (ref_oracle.cpp), hfscf's SCF loop
    (oracle/scf_ref.py). About 22 GB of stored ERIs; the analytic gradient takes hours of CPU."""
    from oracle import basis_ref, oneint_ref, scf_ref

    Z, X = basis_ref.read_xyz(os.path.join(HERE, "h2o_16.xyz"))
    sh, pure, n = basis_ref.shells_for(Z, "cc-pvdz")
    t = time.time()
    S, T, V = oneint_ref.oneint(sh, pure, Z, X)
    Hc = T + V
    enuc = oneint_ref.enuc(Z, X)
    print("h2o16 oneint", time.time() - t, flush=True)
    orc = oracle.Oracle(sh, X, pure, 1e-14, screen=True)
    t = time.time()
    orc.build_eri()
    print("h2o16 eri build", time.time() - t, flush=True)
    r = scf_ref.rhf(S, Hc, int(sum(Z)) // 2, orc.jk, diis_range=8, energy_tol=1e-12, rms_tol=1e-11)
    D = r["D"]
    J, K = orc.jk(D)
    F = Hc + 2.0 * J - K  # Fock matrix of the converged density
    energy = r["energy_elec"] + enuc
    energy_final = float(np.sum(D * (Hc + F))) + enuc
    print("h2o16 energy", energy, energy_final, r["iterations"], r["converged"], flush=True)
    # the energies first (the analytic gradient below takes hours of CPU)
    np.savez_compressed(os.path.join(HERE, "energy_h2o16.npz"), energy=energy, energy_final=energy_final,
                        iterations=r["iterations"])
    del orc  # frees the stored ERIs before the on-the-fly derivative integrals
    orc = oracle.Oracle(sh, X, pure, 1e-14, screen=True)
    t = time.time()
    g1 = oneint_ref.oneint_gradient(sh, pure, Z, X, D, D @ F @ D) + oneint_ref.enuc_gradient(Z, X)
    print("h2o16 one-electron gradient", time.time() - t, flush=True)
    t = time.time()
    g2 = orc.grad2e(D)
    print("h2o16 two-electron gradient", time.time() - t, np.abs((g1 + g2).sum(0)).max(), flush=True)
    np.savez_compressed(os.path.join(HERE, "grad_h2o16.npz"), energy=energy, energy_final=energy_final,
                        iterations=r["iterations"], g=g1 + g2, g2e=g2)


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
