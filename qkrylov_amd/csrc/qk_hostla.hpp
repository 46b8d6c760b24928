// Small dense eigenproblems solved on the host in FP64: the Lanczos
// tridiagonal (implicit-shift QL with a Wilkinson-type shift and 1e-14
// deflation, at most 1000 sweeps; the sweep rules of the reference
// src/solvers/lanczos.cpp:25-98 and ftlm.cpp:61-119, so that the
// eigenvalues, and hence the Lanczos stopping iteration, are the same) and
// Hermitian matrices (the Davidson projected matrix, which the reference
// diagonalises with Jacobi rotations, davidson.cpp:22-82, and the
// block-Lanczos matrix): Householder tridiagonalisation followed by the same
// QL.  Any converged solver gives the same eigenvalues.
#pragma once

#include "qk_common.hpp"

#include <vector>

namespace qk {

// Eigenvalues of the symmetric tridiagonal (diag a[0..n), off-diag b[0..n-1)).
// When z0 is non-null it receives the first component of every eigenvector;
// when Z is non-null it receives the full n*n eigenvector matrix (row-major,
// column i = eigenvector i).
void tridiag_ql(const double* a, const double* b, int n, double* eig, double* z0, double* Z);

// Ground state of the tridiagonal: lowest eigenvalue and its eigenvector.
double tridiag_ground(const std::vector<double>& a, const std::vector<double>& b, int n,
                      std::vector<double>* vec);

// Eigen-decomposition of a small Hermitian matrix (row-major n*n), ascending
// eigenvalues; vecs[k*n + j] = component j of eigenvector k.  Householder
// tridiagonalisation and tridiag_ql: O(n^3) with a small constant, for the
// solvers' per-iteration projected matrices.
void hermitian_eigh(const std::vector<cplx>& A, int n, std::vector<double>& evals,
                    std::vector<cplx>& vecs);

}  // namespace qk
