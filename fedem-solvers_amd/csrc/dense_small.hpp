// Small dense host kernels for the reducer's bookkeeping matrices (block
// Gram matrices of the Lanczos process and its projected eigenproblem).
// Sizes are at most a few hundred; everything is row-major.
#pragma once
#include <vector>

namespace fgpu {

// Cholesky A = R^T R of an n x n SPD matrix; returns R (upper) or false.
bool chol_upper(int n, const std::vector<double>& A, std::vector<double>& R);
// Inverse of an upper-triangular n x n matrix.
void inv_upper(int n, const std::vector<double>& R, std::vector<double>& Ri);
// Symmetric eigen-decomposition: A = V diag(w) V^T, w ascending.
// Householder tridiagonalisation followed by implicit QL with Wilkinson-type
// shifts (Golub & Van Loan, "Matrix Computations", sections 8.3 and 8.3.5).
void sym_eig(int n, const std::vector<double>& A, std::vector<double>& w, std::vector<double>& V);
// All eigenvalues w (ascending) and the eigenvectors of the k largest only,
// V (n x k row-major, column c belongs to w[n-1-c]): Householder
// tridiagonalisation with stored reflectors, implicit QL for the
// eigenvalues, inverse iteration on the tridiagonal (clusters kept
// orthogonal), back-transformation by the reflectors.
void sym_eig_top(int n, const std::vector<double>& A, int k, std::vector<double>& w, std::vector<double>& V);

// Lowest nval eigenvalues (ascending) of the generalized problem A z = l B z
// with A symmetric and B symmetric positive definite (the task of LAPACK
// DSYGVX with ITYPE 1, RANGE 'I', 1..nval): B = R^T R, C = R^-T A R^-1,
// eigenvalues of C by sym_eig.  Both matrices n x n, only their upper
// triangles are read.  Returns 0, or n + j when the leading minor of order j
// of B is not positive definite (DSYGVX's INFO convention).
int sym_gen_eig_lowest(int n, const double* A, const double* B, int nval, double* lambda);

}  // namespace fgpu
