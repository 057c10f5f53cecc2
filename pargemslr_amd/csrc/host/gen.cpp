#include "gen.hpp"

#include <cmath>

namespace pgs {

template <typename T>
Csr<T> laplacian(int nx, int ny, int nz, T shift, double ax, double ay, double az) {
  const int64_t n64 = (int64_t)nx * ny * nz;
  Csr<T> A;
  A.nrows = A.ncols = (int)n64;
  A.ia.assign(A.nrows + 1, 0);
  if (n64 == 0) return A;
  int ones = (nx == 1) + (ny == 1) + (nz == 1);
  const T vd = T(ones > 0 ? 4.0 : 6.0) - shift;
  A.ja.reserve(n64 * (ones > 0 ? 5 : 7));
  A.a.reserve(n64 * (ones > 0 ? 5 : 7));
  const int nxy = nx * ny;
  for (int ii = 0; ii < A.nrows; ++ii) {
    int k = ii / nxy, i = (ii - k * nxy) / nx, j = ii - k * nxy - i * nx;
    if (k > 0) { A.ja.push_back(ii - nxy); A.a.push_back(T(-1.0 + az)); }
    if (i > 0) { A.ja.push_back(ii - nx); A.a.push_back(T(-1.0 + ay)); }
    if (j > 0) { A.ja.push_back(ii - 1); A.a.push_back(T(-1.0 + ax)); }
    A.ja.push_back(ii); A.a.push_back(vd);
    if (j < nx - 1) { A.ja.push_back(ii + 1); A.a.push_back(T(-1.0 - ax)); }
    if (i < ny - 1) { A.ja.push_back(ii + nx); A.a.push_back(T(-1.0 - ay)); }
    if (k < nz - 1) { A.ja.push_back(ii + nxy); A.a.push_back(T(-1.0 - az)); }
    A.ia[ii + 1] = (int)A.ja.size();
  }
  return A;
}

Csr<double> convdiff_upwind(int nx, int ny, int nz, double c) {
  Csr<double> A;
  A.nrows = A.ncols = nx * ny * nz;
  A.ia.assign(A.nrows + 1, 0);
  A.ja.reserve((size_t)A.nrows * 7);
  A.a.reserve((size_t)A.nrows * 7);
  const int nxy = nx * ny;
  const double up = -1.0 - c, down = -1.0, d = 6.0 + 3.0 * c;
  for (int ii = 0; ii < A.nrows; ++ii) {
    int k = ii / nxy, i = (ii - k * nxy) / nx, j = ii - k * nxy - i * nx;
    if (k > 0) { A.ja.push_back(ii - nxy); A.a.push_back(up); }
    if (i > 0) { A.ja.push_back(ii - nx); A.a.push_back(up); }
    if (j > 0) { A.ja.push_back(ii - 1); A.a.push_back(up); }
    A.ja.push_back(ii); A.a.push_back(d);
    if (j < nx - 1) { A.ja.push_back(ii + 1); A.a.push_back(down); }
    if (i < ny - 1) { A.ja.push_back(ii + nx); A.a.push_back(down); }
    if (k < nz - 1) { A.ja.push_back(ii + nxy); A.a.push_back(down); }
    A.ia[ii + 1] = (int)A.ja.size();
  }
  return A;
}

uint64_t rng_u64(uint64_t seed, uint64_t idx) {
  uint64_t z = seed * 0xD1B54A32D192ED03ULL + (idx + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void rng_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double u = (double)(rng_u64(seed, (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
    out[i] = lo + (hi - lo) * u;
  }
}

void rng_normal(uint64_t seed, int64_t n, double* out) {
  const double two_pi = 6.283185307179586;
  for (int64_t i = 0; i < n; ++i) {
    double u1 = ((double)(rng_u64(seed, 2 * (uint64_t)i) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = (double)(rng_u64(seed, 2 * (uint64_t)i + 1) >> 11) * (1.0 / 9007199254740992.0);
    out[i] = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2);
  }
}

template Csr<double> laplacian<double>(int, int, int, double, double, double, double);
template Csr<zdouble> laplacian<zdouble>(int, int, int, zdouble, double, double, double);

}  // namespace pgs
