// Small dense FP64 linear algebra for the host SCF loop (no Eigen/LAPACK in this image).
//   eigh:      symmetric eigensolver, Householder tridiagonalisation followed by implicit
//              QL iterations with Wilkinson shifts (Golub & Van Loan, Matrix Computations,
//              4th ed., sections 8.3.1 and 8.3.3); eigenvalues ascending.
//   qr_solve:  square solve by Householder QR (as hfscf's DIIS does with Eigen's
//              householderQr, hfscf: src/scfsolver/hfscf_diis.cpp:100).
// Matrices are row-major.
#pragma once
#include <vector>

namespace hfscf {

// A (n x n, symmetric) -> eigenvalues w (ascending) and eigenvectors as columns of V.
void eigh(int n, const double* A, double* w, double* V);

// Solve A x = b for square A (n x n) with Householder QR.
std::vector<double> qr_solve(int n, std::vector<double> A, std::vector<double> b);

// C = op(A) * op(B), all n x n.
void gemm(int n, const double* A, bool ta, const double* B, bool tb, double* C);

}  // namespace hfscf
