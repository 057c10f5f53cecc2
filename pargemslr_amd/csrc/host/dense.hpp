// Small dense kernels for the low-rank setup: complex Schur decomposition with
// eigenvalue reordering, dense inverse, orthonormalization.  Column-major
// storage, sizes up to a few hundred.  The reference delegates these to LAPACK
// (?gehrd/?hseqr/?trsen/?getrf/?getri via DenseMatrixClass, ParGeMSLR/SRC/
// matrices/matrixops.cpp:726-3997); LAPACK is not a dependency here, so the
// textbook algorithms are implemented directly: Householder reduction to
// Hessenberg form, single-shift (Wilkinson) implicit QR, adjacent swaps of
// diagonal entries by Givens rotations, LU with partial pivoting.
#pragma once

#include <complex>
#include <vector>

namespace pgs {
namespace dense {

using cplx = std::complex<double>;

// A (n x n, column-major, overwritten by upper-triangular T), Q (n x n) such
// that A_in = Q T Q^H.  Returns 0 on success, 1 if QR did not converge.
int schur(int n, std::vector<cplx>& A, std::vector<cplx>& Q);

// Reorders the Schur form so that the eigenvalues flagged in `select` come
// first (stable within selected / unselected).  Updates T and Q.
void schur_reorder(int n, std::vector<cplx>& T, std::vector<cplx>& Q, const std::vector<char>& select);

// Unit-norm eigenvectors of upper-triangular T (column i for T(i,i)).
void tri_eigvecs(int n, const std::vector<cplx>& T, std::vector<cplx>& Y);

// In-place inverse by LU with partial pivoting.  Returns 0, or 1 if singular.
template <typename S>
int invert(int n, std::vector<S>& A);

// Orthonormalizes the columns of M (rows x cols) with two passes of Gram-
// Schmidt, dropping columns whose norm falls below tol * (their input norm);
// returns the number of columns kept (packed at the front).
template <typename S>
int orthonormalize(int rows, int cols, std::vector<S>& M, double tol);

}  // namespace dense
}  // namespace pgs
