// C ABI of the problem generators and deterministic random vectors.
#include <cstdlib>
#include <cstring>

#include "../../../include/pargemslr_amd.h"
#include "gen.hpp"

using namespace pgs;

namespace {
template <typename T>
int export_csr(const Csr<T>& A, int* n, int64_t* nnz, int** ia, int** ja, double** a) {
  *n = A.nrows;
  *nnz = A.nnz();
  *ia = (int*)std::malloc(sizeof(int) * (A.nrows + 1));
  *ja = (int*)std::malloc(sizeof(int) * std::max<int64_t>(1, A.nnz()));
  *a = (double*)std::malloc(sizeof(T) * std::max<int64_t>(1, A.nnz()));
  if (!*ia || !*ja || !*a) {
    std::free(*ia);
    std::free(*ja);
    std::free(*a);
    return PGS_ERR_OOM;
  }
  std::memcpy(*ia, A.ia.data(), sizeof(int) * (A.nrows + 1));
  std::memcpy(*ja, A.ja.data(), sizeof(int) * A.nnz());
  std::memcpy(*a, A.a.data(), sizeof(T) * A.nnz());
  return 0;
}
}  // namespace

extern "C" {

int pgs_gen_laplacian(int nx, int ny, int nz, double shift_re, double shift_im, double ax, double ay, double az,
                      int is_complex, int* n, int64_t* nnz, int** ia, int** ja, double** a) {
  if (nx < 1 || ny < 1 || nz < 1 || !n || !nnz || !ia || !ja || !a) return PGS_ERROR_INVALID_PARAM;
  if ((int64_t)nx * ny * nz > 2000000000LL) return PGS_ERROR_INVALID_PARAM;
  if (!is_complex) return export_csr(laplacian<double>(nx, ny, nz, shift_re, ax, ay, az), n, nnz, ia, ja, a);
  return export_csr(laplacian<zdouble>(nx, ny, nz, zdouble(shift_re, shift_im), ax, ay, az), n, nnz, ia, ja, a);
}

int pgs_gen_convdiff(int nx, int ny, int nz, double c, int* n, int64_t* nnz, int** ia, int** ja, double** a) {
  if (nx < 1 || ny < 1 || nz < 1 || !n || !nnz || !ia || !ja || !a) return PGS_ERROR_INVALID_PARAM;
  return export_csr(convdiff_upwind(nx, ny, nz, c), n, nnz, ia, ja, a);
}

void pgs_free(void* p) { std::free(p); }

int pgs_rng_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out) {
  if (n < 0 || (n > 0 && !out)) return PGS_ERROR_INVALID_PARAM;
  rng_uniform(seed, n, lo, hi, out);
  return 0;
}

int pgs_rng_normal(uint64_t seed, int64_t n, double* out) {
  if (n < 0 || (n > 0 && !out)) return PGS_ERROR_INVALID_PARAM;
  rng_normal(seed, n, out);
  return 0;
}

}  // extern "C"
