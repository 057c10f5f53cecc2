// Host-side sparse data model used by the GeMSLR setup.
//
// CSR with int32 row pointers / column indices, values of type T (double or
// std::complex<double>).  Mirrors the data model of CsrMatrixClass
// (ParGeMSLR/SRC/matrices/csr_matrix.hpp:26-716): 0-based CSR, square or
// rectangular, rows not necessarily sorted.
#pragma once

#include <algorithm>
#include <atomic>
#include <complex>
#include <cstdint>
#include <cmath>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

namespace pgs {

using zdouble = std::complex<double>;

template <typename T> struct is_complex : std::false_type {};
template <typename R> struct is_complex<std::complex<R>> : std::true_type {};

inline double sabs(double v) { return std::fabs(v); }
inline double sabs(const zdouble& v) { return std::abs(v); }
inline double sconj(double v) { return v; }
inline zdouble sconj(const zdouble& v) { return std::conj(v); }
inline double sreal(double v) { return v; }
inline double sreal(const zdouble& v) { return v.real(); }

template <typename T>
struct Csr {
  int nrows = 0, ncols = 0;
  std::vector<int> ia{0}, ja;
  std::vector<T> a;
  int64_t nnz() const { return (int64_t)ja.size(); }
};

// Runs f(i) for i in [0, n) over a small pool of std::threads (dynamic chunks).
// The setup code uses this instead of OpenMP so that the library does not pull
// a second OpenMP runtime into processes that already host one.
inline void parallel_for(int n, const std::function<void(int)>& f, int nthreads = 0) {
  if (nthreads <= 0) {
    nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads <= 0) nthreads = 1;
    if (nthreads > 16) nthreads = 16;
  }
  if (n <= 1 || nthreads == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  std::atomic<int> next{0};
  // the first exception a worker throws (std::bad_alloc from a body's
  // allocation, say) stops the remaining chunks and is rethrown on the
  // calling thread after every worker has joined
  std::exception_ptr first;
  std::mutex first_m;
  int nt = std::min(nthreads, n);
  int chunk = std::max(1, n / (nt * 32));
  for (int t = 0; t < nt; ++t) {
    pool.emplace_back([&]() {
      try {
        for (;;) {
          int i0 = next.fetch_add(chunk);
          if (i0 >= n) break;
          int i1 = std::min(n, i0 + chunk);
          for (int i = i0; i < i1; ++i) f(i);
        }
      } catch (...) {
        next.store(n);
        std::lock_guard<std::mutex> g(first_m);
        if (!first) first = std::current_exception();
      }
    });
  }
  for (auto& th : pool) th.join();
  if (first) std::rethrow_exception(first);
}

template <typename T>
Csr<T> csr_transpose(const Csr<T>& A) {
  Csr<T> B;
  B.nrows = A.ncols;
  B.ncols = A.nrows;
  B.ia.assign(B.nrows + 1, 0);
  for (int64_t k = 0; k < A.nnz(); ++k) B.ia[A.ja[k] + 1]++;
  for (int i = 0; i < B.nrows; ++i) B.ia[i + 1] += B.ia[i];
  B.ja.resize(A.nnz());
  B.a.resize(A.nnz());
  std::vector<int> pos(B.ia.begin(), B.ia.end() - 1);
  for (int i = 0; i < A.nrows; ++i)
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k) {
      int p = pos[A.ja[k]]++;
      B.ja[p] = i;
      B.a[p] = A.a[k];
    }
  return B;
}

// Sorts the entries of every row by column index.
template <typename T>
void csr_sort_rows(Csr<T>& A) {
  std::vector<std::pair<int, T>> buf;
  for (int i = 0; i < A.nrows; ++i) {
    int k1 = A.ia[i], k2 = A.ia[i + 1];
    bool sorted = true;
    for (int k = k1 + 1; k < k2; ++k)
      if (A.ja[k] < A.ja[k - 1]) { sorted = false; break; }
    if (sorted) continue;
    buf.clear();
    for (int k = k1; k < k2; ++k) buf.emplace_back(A.ja[k], A.a[k]);
    std::sort(buf.begin(), buf.end(), [](const std::pair<int, T>& x, const std::pair<int, T>& y) { return x.first < y.first; });
    for (int k = k1; k < k2; ++k) { A.ja[k] = buf[k - k1].first; A.a[k] = buf[k - k1].second; }
  }
}

// Sub-matrix A(rows, cols) with rows/cols given as index lists (any order);
// the result's row r is A's row rows[r], columns are renumbered by position in
// cols; entries whose column is not in cols are skipped.  Entry order within
// a row follows A.
template <typename T>
Csr<T> csr_submatrix(const Csr<T>& A, const std::vector<int>& rows, const std::vector<int>& cols,
                     std::vector<int>* colmap_buf = nullptr) {
  std::vector<int> local;
  std::vector<int>& cmap = colmap_buf ? *colmap_buf : local;
  if ((int)cmap.size() < A.ncols) cmap.assign(A.ncols, -1);
  for (int c = 0; c < (int)cols.size(); ++c) cmap[cols[c]] = c;
  Csr<T> B;
  B.nrows = (int)rows.size();
  B.ncols = (int)cols.size();
  B.ia.assign(B.nrows + 1, 0);
  for (int r = 0; r < B.nrows; ++r) {
    int i = rows[r], cnt = 0;
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k)
      if (cmap[A.ja[k]] >= 0) ++cnt;
    B.ia[r + 1] = B.ia[r] + cnt;
  }
  B.ja.resize(B.ia[B.nrows]);
  B.a.resize(B.ia[B.nrows]);
  for (int r = 0; r < B.nrows; ++r) {
    int i = rows[r], p = B.ia[r];
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k) {
      int c = cmap[A.ja[k]];
      if (c >= 0) { B.ja[p] = c; B.a[p] = A.a[k]; ++p; }
    }
  }
  for (int c = 0; c < (int)cols.size(); ++c) cmap[cols[c]] = -1;
  return B;
}

// Contiguous block A[r0:r0+nr, c0:c0+nc] with local column numbering.
template <typename T>
Csr<T> csr_block(const Csr<T>& A, int r0, int nr, int c0, int nc) {
  Csr<T> B;
  B.nrows = nr;
  B.ncols = nc;
  B.ia.assign(nr + 1, 0);
  for (int r = 0; r < nr; ++r) {
    int cnt = 0;
    for (int k = A.ia[r0 + r]; k < A.ia[r0 + r + 1]; ++k) {
      int c = A.ja[k];
      if (c >= c0 && c < c0 + nc) ++cnt;
    }
    B.ia[r + 1] = B.ia[r] + cnt;
  }
  B.ja.resize(B.ia[nr]);
  B.a.resize(B.ia[nr]);
  for (int r = 0; r < nr; ++r) {
    int p = B.ia[r];
    for (int k = A.ia[r0 + r]; k < A.ia[r0 + r + 1]; ++k) {
      int c = A.ja[k];
      if (c >= c0 && c < c0 + nc) { B.ja[p] = c - c0; B.a[p] = A.a[k]; ++p; }
    }
  }
  return B;
}

// Symmetric permutation B = A(p, p): B(i, j) = A(p[i], p[j]).
template <typename T>
Csr<T> csr_sym_permute(const Csr<T>& A, const std::vector<int>& p) {
  std::vector<int> rp(A.nrows);
  for (int i = 0; i < A.nrows; ++i) rp[p[i]] = i;
  Csr<T> B;
  B.nrows = A.nrows;
  B.ncols = A.ncols;
  B.ia.assign(B.nrows + 1, 0);
  for (int i = 0; i < B.nrows; ++i) B.ia[i + 1] = B.ia[i] + (A.ia[p[i] + 1] - A.ia[p[i]]);
  B.ja.resize(A.nnz());
  B.a.resize(A.nnz());
  parallel_for(B.nrows, [&](int i) {
    int q = B.ia[i];
    for (int k = A.ia[p[i]]; k < A.ia[p[i] + 1]; ++k, ++q) { B.ja[q] = rp[A.ja[k]]; B.a[q] = A.a[k]; }
  });
  return B;
}

// y = alpha*A*x + beta*y  (host, used by the setup's Arnoldi operator).
template <typename T>
void csr_matvec(const Csr<T>& A, T alpha, const T* x, T beta, T* y) {
  for (int i = 0; i < A.nrows; ++i) {
    T s = T(0);
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k) s += A.a[k] * x[A.ja[k]];
    y[i] = (beta == T(0) ? T(0) : beta * y[i]) + alpha * s;
  }
}

}  // namespace pgs
