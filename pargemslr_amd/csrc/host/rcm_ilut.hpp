// Local reordering (RCM) and ILUT factorization for the GeMSLR diagonal blocks.
#pragma once

#include "csr.hpp"

namespace pgs {

// Reverse Cuthill-McKee order of the pattern of A + A^T.  perm[new] = old.
// Follows CsrMatrixRcmHost (ParGeMSLR/SRC/matrices/matrixops.cpp:8533-8641):
// per connected component, the root is the minimum-degree vertex of the first
// unnumbered component, moved to a pseudo-peripheral vertex by repeated BFS
// (matrixops.cpp:8700-8790); vertices are numbered in BFS order and the
// component's numbering is reversed.  The reference's per-node neighbour sort
// at matrixops.cpp:8760 sorts an empty range, so neighbours keep adjacency
// order; this restatement does the same.
template <typename T>
std::vector<int> rcm_order(const Csr<T>& A);

// ILUT factor of one diagonal block: A ~= L (D + U), L unit lower (strict part
// stored), U strictly upper, Dinv = 1 / pivots.  Row entries are stored sorted
// by column.
template <typename T>
struct IluFactor {
  int n = 0;
  Csr<T> L, U;
  std::vector<T> Dinv;
  int64_t nnz() const { return (int64_t)n + L.nnz() + U.nnz(); }
};

// Dual-threshold ILUT following IluClass::SetupILUT
// (ParGeMSLR/SRC/preconditioners/ilu.cpp:1647-2180): row-wise IKJ elimination
// in increasing column order, a fill-in is dropped when |l*u| < tol*||a_i||_1/nnz_i,
// each of L and U keeps the `lfil` largest entries by magnitude, pivots with
// |d| < 1e-6 are replaced by 1e-6.  Returns 0 or 103 (empty row).
template <typename T>
int ilut_factor(const Csr<T>& A, double droptol, int lfil, IluFactor<T>& F);

}  // namespace pgs
