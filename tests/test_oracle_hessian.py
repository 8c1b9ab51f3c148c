"""Pins of the oracle's second-derivative ERIs and two-electron Hessian (CPU, no GPU).

* Shell quartets: each of the 45 second-derivative blocks (coordinate pairs over the centres A,
  B, C) against a central finite difference of the oracle's first-derivative blocks under a shift
  of one shell's centre.
* Hessian: hfscf's Coulomb and exchange parts at a fixed density against a fourth-order central
  difference of the oracle's Coulomb and exchange gradients (grad2e_parts) under atom
  displacements; symmetry and translational invariance (every row sums to zero over the atoms
  for each direction).
"""
import numpy as np
import pytest

import oracle
from conftest import golden
from oracle import basis_ref


def _moved(sh, X, shell, d, h):
    """Shells and geometry with only `shell`'s centre moved by h along d (on a duplicate atom)."""
    sh2 = {k: v.copy() for k, v in sh.items()}
    atom = int(sh["atom"][shell])
    sh2["atom"][shell] = len(X)
    X2 = np.vstack([X, X[atom]])
    X2[len(X), d] += h
    return sh2, X2


@pytest.mark.parametrize("q", [(5, 2, 4, 0), (5, 3, 7, 1), (2, 3, 5, 8), (5, 4, 3, 2)])
def test_deriv2_blocks_against_finite_differences(q):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    sh, _, _ = basis_ref.shells_for(Z, "6-31g*")  # Cartesian, s p d
    o = oracle.Oracle(sh, X, False, 1e-14, screen=False)
    der2 = o.shell_quartet_deriv2(*q)
    h = 1e-5
    pidx = 0
    worst = 0.0
    for X1 in range(9):
        for X2 in range(X1, 9):
            c1, d1 = divmod(X1, 3)
            shp, Xp = _moved(sh, X, q[c1], d1, h)
            shm, Xm = _moved(sh, X, q[c1], d1, -h)
            fd = (oracle.Oracle(shp, Xp, False, 1e-14, screen=False).shell_quartet_deriv1(*q)[X2] -
                  oracle.Oracle(shm, Xm, False, 1e-14, screen=False).shell_quartet_deriv1(*q)[X2]) / (2 * h)
            err = np.abs(fd - der2[pidx]).max()
            worst = max(worst, err)
            assert err < 2e-7 * max(1.0, np.abs(fd).max()), (q, X1, X2, err)
            pidx += 1
    assert worst > 0.0 or np.abs(der2).max() == 0.0


def _density(n, nocc, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, nocc)) / np.sqrt(n)
    return C @ C.T


@pytest.mark.parametrize("basis", ["sto-3g", "6-31g*", "cc-pvdz"])
def test_hess2e_against_gradient_differences(basis):
    Z, X = basis_ref.read_xyz(golden("h2o.xyz"))
    sh, pure, nbf = basis_ref.shells_for(Z, basis)
    D = _density(nbf, 5, 7)
    hc, hx = oracle.Oracle(sh, X, pure, 1e-14, screen=False).hess2e(D)
    n3 = 3 * len(X)
    fc, fx = np.zeros((n3, n3)), np.zeros((n3, n3))
    h = 2e-3
    for col in range(n3):
        a, d = divmod(col, 3)
        parts = []
        for step in (2, 1, -1, -2):
            Xs = X.copy()
            Xs[a, d] += step * h
            gc, gx = oracle.Oracle(sh, Xs, pure, 1e-14, screen=False).grad2e_parts(D)
            parts.append((gc.reshape(-1), gx.reshape(-1)))
        for k, out in ((0, fc), (1, fx)):
            out[:, col] = (-parts[0][k] + 8 * parts[1][k] - 8 * parts[2][k] + parts[3][k]) / (12 * h)
    scale = max(np.abs(hc).max(), np.abs(hx).max())
    assert np.abs(hc - fc).max() < 1e-7 * scale, np.abs(hc - fc).max()
    assert np.abs(hx - fx).max() < 1e-7 * scale, np.abs(hx - fx).max()
    # symmetry and translational invariance
    for H in (hc, hx):
        assert np.abs(H - H.T).max() < 1e-11 * scale
        assert np.abs(H.reshape(n3, -1, 3).sum(axis=1)).max() < 1e-10 * scale
    # the exchange part is not a multiple of the Coulomb part (both are exercised)
    assert np.abs(hx).max() > 1e-3 * np.abs(hc).max()
