#include "dense.hpp"

#include <cfloat>
#include <cmath>

namespace pgs {
namespace dense {

namespace {

inline double cabs1(const cplx& z) { return std::abs(z.real()) + std::abs(z.imag()); }
inline double cj(double x) { return x; }
inline cplx cj(const cplx& z) { return std::conj(z); }

// Plane rotation G = [c s; -conj(s) c] with G [f; g] = [r; 0], c real.
inline void lartg(const cplx& f, const cplx& g, double& c, cplx& s) {
  double af = std::abs(f), ag = std::abs(g);
  if (ag == 0.0) { c = 1.0; s = 0.0; return; }
  if (af == 0.0) { c = 0.0; s = std::conj(g) / ag; return; }
  double nrm = std::hypot(af, ag);
  c = af / nrm;
  s = (f / af) * std::conj(g) / nrm;
}

// rows k, k+1 <- G * rows, for columns [c0, c1)
inline void rot_rows(int n, std::vector<cplx>& A, int k, int c0, int c1, double c, const cplx& s) {
  for (int j = c0; j < c1; ++j) {
    cplx t1 = A[k + (size_t)j * n], t2 = A[k + 1 + (size_t)j * n];
    A[k + (size_t)j * n] = c * t1 + s * t2;
    A[k + 1 + (size_t)j * n] = -std::conj(s) * t1 + c * t2;
  }
}

// cols k, k+1 <- cols * G^H, for rows [r0, r1)
inline void rot_cols(int n, std::vector<cplx>& A, int k, int r0, int r1, double c, const cplx& s) {
  cplx* a0 = &A[(size_t)k * n];
  cplx* a1 = &A[(size_t)(k + 1) * n];
  for (int i = r0; i < r1; ++i) {
    cplx t1 = a0[i], t2 = a1[i];
    a0[i] = c * t1 + std::conj(s) * t2;
    a1[i] = -s * t1 + c * t2;
  }
}

}  // namespace

int schur(int n, std::vector<cplx>& A, std::vector<cplx>& Q) {
  Q.assign((size_t)n * n, cplx(0));
  for (int i = 0; i < n; ++i) Q[i + (size_t)i * n] = 1.0;
  if (n <= 1) return 0;
  auto at = [&](int i, int j) -> cplx& { return A[i + (size_t)j * n]; };
  // Householder reduction to upper Hessenberg form
  std::vector<cplx> v(n), tmp(n);
  for (int k = 0; k + 2 < n; ++k) {
    double xn = 0.0;
    for (int i = k + 1; i < n; ++i) xn += std::norm(at(i, k));
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    cplx x0 = at(k + 1, k);
    cplx phase = (std::abs(x0) > 0.0) ? x0 / std::abs(x0) : cplx(1.0);
    cplx alpha = -phase * xn;
    for (int i = k + 1; i < n; ++i) v[i] = at(i, k);
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += std::norm(v[i]);
    vn = std::sqrt(vn);
    if (vn == 0.0) continue;
    for (int i = k + 1; i < n; ++i) v[i] /= vn;
    // A <- H A  (rows k+1..n-1), H = I - 2 v v^H
    for (int j = k; j < n; ++j) {
      cplx d = 0.0;
      for (int i = k + 1; i < n; ++i) d += std::conj(v[i]) * at(i, j);
      d *= 2.0;
      for (int i = k + 1; i < n; ++i) at(i, j) -= v[i] * d;
    }
    // A <- A H, Q <- Q H (columns k+1..n-1)
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<cplx>& M = pass == 0 ? A : Q;
      for (int i = 0; i < n; ++i) {
        cplx d = 0.0;
        for (int j = k + 1; j < n; ++j) d += M[i + (size_t)j * n] * v[j];
        d *= 2.0;
        for (int j = k + 1; j < n; ++j) M[i + (size_t)j * n] -= d * std::conj(v[j]);
      }
    }
    at(k + 1, k) = alpha;
    for (int i = k + 2; i < n; ++i) at(i, k) = 0.0;
  }
  // shifted QR on the Hessenberg matrix
  const double eps = DBL_EPSILON;
  double anorm = 0.0;
  for (size_t t = 0; t < A.size(); ++t) anorm = std::max(anorm, std::abs(A[t]));
  const double small = std::max(anorm, 1e-300) * eps;
  int hi = n - 1, iter = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; --l) {
      double s = cabs1(at(l - 1, l - 1)) + cabs1(at(l, l));
      if (s == 0.0) s = anorm;
      if (cabs1(at(l, l - 1)) <= eps * s || cabs1(at(l, l - 1)) <= small) { at(l, l - 1) = 0.0; break; }
    }
    if (l == hi) { --hi; iter = 0; continue; }
    if (++total > 60 * n) return 1;
    ++iter;
    cplx mu;
    if (iter % 11 == 10) {
      mu = at(hi, hi) + cplx(1.5 * std::abs(at(hi, hi - 1)), 0.0);
    } else {
      cplx a = at(hi - 1, hi - 1), b = at(hi - 1, hi), c = at(hi, hi - 1), d = at(hi, hi);
      cplx tr = a + d, det = a * d - b * c;
      cplx disc = std::sqrt(tr * tr - 4.0 * det);
      cplx m1 = 0.5 * (tr + disc), m2 = 0.5 * (tr - disc);
      mu = (std::abs(m1 - d) < std::abs(m2 - d)) ? m1 : m2;
    }
    for (int k = l; k < hi; ++k) {
      cplx f, g;
      if (k == l) { f = at(l, l) - mu; g = at(l + 1, l); }
      else { f = at(k, k - 1); g = at(k + 1, k - 1); }
      double c;
      cplx s;
      lartg(f, g, c, s);
      rot_rows(n, A, k, (k == l) ? l : k - 1, n, c, s);
      if (k > l) at(k + 1, k - 1) = 0.0;
      rot_cols(n, A, k, 0, std::min(k + 3, hi + 1), c, s);
      rot_cols(n, Q, k, 0, n, c, s);
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) at(i, j) = 0.0;
  return 0;
}

void schur_reorder(int n, std::vector<cplx>& T, std::vector<cplx>& Q, const std::vector<char>& select) {
  std::vector<char> sel(select.begin(), select.end());
  int placed = 0;
  for (int j = 0; j < n; ++j) {
    if (!sel[j]) continue;
    for (int k = j - 1; k >= placed; --k) {
      // swap diagonal entries k and k+1
      cplx a = T[k + (size_t)k * n], b = T[k + 1 + (size_t)(k + 1) * n];
      cplx f = T[k + (size_t)(k + 1) * n], g = b - a;
      double c;
      cplx s;
      lartg(f, g, c, s);
      rot_rows(n, T, k, k, n, c, s);
      rot_cols(n, T, k, 0, k + 2, c, s);
      rot_cols(n, Q, k, 0, n, c, s);
      T[k + 1 + (size_t)k * n] = 0.0;
      std::swap(sel[k], sel[k + 1]);
    }
    ++placed;
  }
}

void tri_eigvecs(int n, const std::vector<cplx>& T, std::vector<cplx>& Y) {
  Y.assign((size_t)n * n, cplx(0));
  double tn = 0.0;
  for (const cplx& z : T) tn = std::max(tn, std::abs(z));
  const double smallnum = std::max(tn, 1e-300) * DBL_EPSILON;
  for (int i = 0; i < n; ++i) {
    cplx* y = &Y[(size_t)i * n];
    y[i] = 1.0;
    cplx lam = T[i + (size_t)i * n];
    for (int j = i - 1; j >= 0; --j) {
      cplx s = 0.0;
      for (int l = j + 1; l <= i; ++l) s += T[j + (size_t)l * n] * y[l];
      cplx den = T[j + (size_t)j * n] - lam;
      if (std::abs(den) < smallnum) den = smallnum;
      y[j] = -s / den;
    }
    double nr = 0.0;
    for (int j = 0; j <= i; ++j) nr += std::norm(y[j]);
    nr = std::sqrt(nr);
    for (int j = 0; j <= i; ++j) y[j] /= nr;
  }
}

template <typename S>
int invert(int n, std::vector<S>& A) {
  std::vector<int> piv(n);
  std::vector<S> inv((size_t)n * n, S(0));
  for (int i = 0; i < n; ++i) inv[i + (size_t)i * n] = S(1);
  auto at = [&](std::vector<S>& M, int i, int j) -> S& { return M[i + (size_t)j * n]; };
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::abs(at(A, k, k));
    for (int i = k + 1; i < n; ++i)
      if (std::abs(at(A, i, k)) > best) { best = std::abs(at(A, i, k)); p = i; }
    if (best == 0.0) return 1;
    if (p != k)
      for (int j = 0; j < n; ++j) { std::swap(at(A, k, j), at(A, p, j)); std::swap(at(inv, k, j), at(inv, p, j)); }
    S d = S(1) / at(A, k, k);
    for (int j = 0; j < n; ++j) { at(A, k, j) *= d; at(inv, k, j) *= d; }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      S f = at(A, i, k);
      if (f == S(0)) continue;
      for (int j = 0; j < n; ++j) { at(A, i, j) -= f * at(A, k, j); at(inv, i, j) -= f * at(inv, k, j); }
    }
  }
  A.swap(inv);
  return 0;
}

template <typename S>
int orthonormalize(int rows, int cols, std::vector<S>& M, double tol) {
  int kept = 0;
  for (int j = 0; j < cols; ++j) {
    S* v = &M[(size_t)j * rows];
    double n0 = 0.0;
    for (int i = 0; i < rows; ++i) n0 += std::norm(v[i]);
    n0 = std::sqrt(n0);
    if (n0 == 0.0) continue;
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < kept; ++q) {
        const S* u = &M[(size_t)q * rows];
        S d = S(0);
        for (int i = 0; i < rows; ++i) d += cj(u[i]) * v[i];
        for (int i = 0; i < rows; ++i) v[i] -= d * u[i];
      }
    double nr = 0.0;
    for (int i = 0; i < rows; ++i) nr += std::norm(v[i]);
    nr = std::sqrt(nr);
    if (nr <= tol * n0) continue;
    S* dst = &M[(size_t)kept * rows];
    for (int i = 0; i < rows; ++i) dst[i] = v[i] / nr;
    ++kept;
  }
  return kept;
}

template int invert<double>(int, std::vector<double>&);
template int invert<cplx>(int, std::vector<cplx>&);
template int orthonormalize<double>(int, int, std::vector<double>&, double);
template int orthonormalize<cplx>(int, int, std::vector<cplx>&, double);

}  // namespace dense
}  // namespace pgs
