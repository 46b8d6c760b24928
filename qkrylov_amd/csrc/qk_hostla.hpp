// Small dense eigenproblems solved on the host in FP64: the Lanczos
// tridiagonal (implicit-shift QL with a Wilkinson-type shift and 1e-14
// deflation, at most 1000 sweeps) and the Davidson projected matrix (cyclic
// largest-element Jacobi, at most 1000 rotations, stop below 1e-15).  The
// sweep and rotation rules are those of the reference (src/solvers/
// lanczos.cpp:25-98, ftlm.cpp:61-119, davidson.cpp:22-82) so that the
// eigenvalues, and hence the Lanczos stopping iteration, are the same.
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
// eigenvalues; vecs[k*n + j] = component j of eigenvector k.
void hermitian_jacobi(const std::vector<cplx>& A, int n, std::vector<double>& evals,
                      std::vector<cplx>& vecs);

}  // namespace qk
