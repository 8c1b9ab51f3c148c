"""ORACLE — test infrastructure only. hfscf's RHF SCF loop semantics in numpy/scipy.

Restates (citations relative to the hfscf tree):
  * symmetric orthogonalisation X = S^-1/2                     src/scfsolver/hfscf_hf_solver.cpp:184-204
  * CORE guess, half density D = C_occ C_occ^T                 :206-242
  * energy E = sum D (H + F) with F the previous extrapolated  :244-252 (H + H at iteration 0)
  * Fock build F = H + 2J - K from a supplied J/K builder      :254-286
  * DIIS: error X^T (FDS - SDF) X, rms = sqrt(mean e^2), evict the max-rms vector before
    appending once more than diis_range are held, start at two vectors, B normalised when
    sum |B| < 1e-14, QR solve                                  src/scfsolver/hfscf_diis.cpp:9-123
  * converged when rms < rms_tol and |dE| < energy_tol        hfscf_hf_solver.cpp:415
Eigen's eigensolver and QR are replaced by scipy.linalg.eigh and numpy.linalg.qr; their
results agree to rounding (the eigenvector sign does not enter D).
"""
import numpy as np
import scipy.linalg as sla


def rhf(S, H, nocc, jk, diis_range=6, energy_tol=1e-10, rms_tol=1e-10, max_iter=100):
    w, U = sla.eigh(S)
    X = (U / np.sqrt(w)) @ U.T

    def density(Fp):
        _, Cp = sla.eigh(Fp)
        C = X @ Cp
        return C[:, :nocc] @ C[:, :nocc].T

    D = density(X.T @ H @ X)
    energies = [float(np.sum(D * (H + H)))]
    e_list, f_list, rms_list = [], [], []
    F = H
    it = 0
    converged = False
    while True:
        it += 1
        J, K = jk(D)
        F = H + 2.0 * J - K
        FDS = F @ D @ S
        err = X.T @ (FDS - FDS.T) @ X
        rms = float(np.sqrt(np.mean(err * err)))
        if len(rms_list) > diis_range:
            pos = int(np.argmax(rms_list))
            del e_list[pos], f_list[pos], rms_list[pos]
        e_list.append(err)
        f_list.append(F.copy())
        rms_list.append(rms)
        if len(rms_list) > 1:
            m = len(e_list)
            B = np.empty((m + 1, m + 1))
            for i in range(m):
                for j in range(m):
                    B[i, j] = np.sum(e_list[i] * e_list[j])
            if np.sum(np.abs(B[:m, :m])) < 1e-14:
                B[:m, :m] /= B[:m, :m].max()
            B[m, :] = -1.0
            B[:, m] = -1.0
            B[m, m] = 0.0
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            Q, R = np.linalg.qr(B)
            c = sla.solve_triangular(R, Q.T @ rhs)
            F = sum(c[i] * f_list[i] for i in range(m))
        D = density(X.T @ F @ X)
        energies.append(float(np.sum(D * (H + F))))
        if rms < rms_tol and abs(energies[-1] - energies[-2]) < energy_tol:
            converged = True
            break
        if it >= max_iter:
            break
    return dict(energy_elec=energies[-1], energies=energies, D=D, F=F, iterations=it, converged=converged)
