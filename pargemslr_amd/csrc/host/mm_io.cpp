// MatrixMarket ingest: coordinate matrices into CSR, coordinate vectors into
// dense arrays.
//
// Rules follow the reference's readers:
//  - banner and size line as mm_read_banner / mm_read_mtx_crd_size
//    (ParGeMSLR/SRC/utils/mmio.cpp:101-222): the first line holds five
//    tokens, the first starting with "%%MatrixMarket" (case-sensitive), the
//    other four compared in lower case; lines starting with '%' are skipped
//    before the "M N nz" line;
//  - matrices as CooMatrixReadFromFile (ParGeMSLR/SRC/matrices/
//    matrixops.cpp:11134-11214 for double, 11335-11453 for complexd): only
//    "matrix coordinate" files; the double reader takes real and integer
//    data and mirrors an off-diagonal entry only for "symmetric" storage; the
//    complex reader also takes complex data and mirrors symmetric (a),
//    hermitian (conj a) and skew-symmetric (-a) entries (a real file read as
//    complex mirrors hermitian as a and skew as -a); the mirrored entry
//    follows its original;
//  - CSR as CooMatrixClass::ToCsr (coo_matrix.cpp:479-512, serial branch of
//    CooMatrixP2CsrMatrixPHost, matrixops.cpp:6168-6300): rows bucketed in
//    COO order, duplicates kept;
//  - vectors as SequentialVectorReadFromFile (ParGeMSLR/SRC/vectors/
//    vectorops.cpp:1116-1180, 1260-1336): length M, zero-filled, entry (I, J)
//    sets v[I] (J ignored, a later entry overwrites); the complex reader
//    requires N == 1.
// Numbers are read with strtol / strtod from a moving position, which is what
// fscanf("%d %d %lg") does (whitespace, including line ends, separates
// fields).  Differences from the reference, all on inputs it does not
// check: an index outside the declared size, a truncated entry list or a
// symmetric expansion beyond 2^31-1 entries returns PGS_ERROR_IO instead of
// writing out of bounds.
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/pargemslr_amd.h"
#include "csr.hpp"

namespace {

struct MmHeader {
  bool coordinate = false;
  char field = 0;     // 'r' real, 'c' complex, 'p' pattern, 'i' integer
  char storage = 0;   // 'g' general, 's' symmetric, 'h' hermitian, 'k' skew-symmetric
  long long M = 0, N = 0, nz = 0;
};

bool read_file(const char* path, std::string& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, got);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

// Parses banner and size line; returns the offset of the first entry or -1.
long long parse_header(const std::string& s, MmHeader& h) {
  size_t pos = 0;
  auto next_line = [&](std::string& line) -> bool {
    if (pos >= s.size()) return false;
    size_t e = s.find('\n', pos);
    if (e == std::string::npos) e = s.size();
    line.assign(s, pos, e - pos);
    pos = e + (e < s.size() ? 1 : 0);
    return true;
  };
  std::string line;
  if (!next_line(line)) return -1;
  std::vector<std::string> tok;
  {
    size_t i = 0;
    while (i < line.size() && tok.size() < 5) {
      while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
      size_t j = i;
      while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
      if (j > i) tok.emplace_back(line, i, j - i);
      i = j;
    }
  }
  if (tok.size() != 5) return -1;
  if (tok[0].compare(0, 14, "%%MatrixMarket") != 0) return -1;
  for (int t = 1; t < 5; ++t)
    for (char& c : tok[t]) c = (char)std::tolower((unsigned char)c);
  if (tok[1] != "matrix") return -1;
  if (tok[2] == "coordinate") h.coordinate = true;
  else if (tok[2] == "array") h.coordinate = false;
  else return -1;
  if (tok[3] == "real") h.field = 'r';
  else if (tok[3] == "complex") h.field = 'c';
  else if (tok[3] == "pattern") h.field = 'p';
  else if (tok[3] == "integer") h.field = 'i';
  else return -1;
  if (tok[4] == "general") h.storage = 'g';
  else if (tok[4] == "symmetric") h.storage = 's';
  else if (tok[4] == "hermitian") h.storage = 'h';
  else if (tok[4] == "skew-symmetric") h.storage = 'k';
  else return -1;
  // size line: skip comments and blank lines
  for (;;) {
    if (!next_line(line)) return -1;
    if (!line.empty() && line[0] == '%') continue;
    const char* p = line.c_str();
    char* e;
    long long v[3];
    int got = 0;
    for (; got < 3; ++got) {
      errno = 0;
      v[got] = std::strtoll(p, &e, 10);
      if (e == p || errno) break;
      p = e;
    }
    if (got == 0) {
      bool blank = true;
      for (char c : line) blank = blank && std::isspace((unsigned char)c);
      if (blank) continue;
    }
    if (got != 3) return -1;
    h.M = v[0];
    h.N = v[1];
    h.nz = v[2];
    break;
  }
  if (h.M < 0 || h.N < 0 || h.nz < 0 || h.M > 2147483647LL || h.N > 2147483647LL) return -1;
  return (long long)pos;
}

// One coordinate entry: two indices and nv values.
struct EntryReader {
  const char* p;
  const char* end;
  bool index(long long& v) {
    char* e;
    errno = 0;
    v = std::strtoll(p, &e, 10);
    if (e == p || errno) return false;
    p = e;
    return true;
  }
  bool value(double& v) {
    char* e;
    v = std::strtod(p, &e);
    if (e == p) return false;
    p = e;
    return true;
  }
};

template <typename T>
int read_matrix(const char* path, int idxin, pgs::Csr<T>& A) {
  constexpr bool cplx = pgs::is_complex<T>::value;
  std::string buf;
  if (!read_file(path, buf)) return PGS_ERROR_IO;
  MmHeader h;
  const long long off = parse_header(buf, h);
  if (off < 0) return PGS_ERROR_IO;
  const bool field_ok = h.field == 'r' || h.field == 'i' || (cplx && h.field == 'c');
  if (!field_ok || !h.coordinate) return PGS_ERROR_IO;
  // the double reader mirrors symmetric storage only
  const bool mirror = h.storage == 's' || (cplx && (h.storage == 'h' || h.storage == 'k'));
  const long long cap = mirror ? 2 * h.nz : h.nz;
  if (cap > 2147483647LL) return PGS_ERROR_IO;
  std::vector<int> ci, cj;
  std::vector<T> cv;
  ci.reserve(cap);
  cj.reserve(cap);
  cv.reserve(cap);
  EntryReader r{buf.c_str() + off, buf.c_str() + buf.size()};
  for (long long k = 0; k < h.nz; ++k) {
    long long I, J;
    double re = 0.0, im = 0.0;
    if (!r.index(I) || !r.index(J) || !r.value(re)) return PGS_ERROR_IO;
    if (h.field == 'c' && !r.value(im)) return PGS_ERROR_IO;
    I -= idxin;
    J -= idxin;
    if (I < 0 || I >= h.M || J < 0 || J >= h.N) return PGS_ERROR_IO;
    T v;
    if constexpr (cplx) v = T(re, im);
    else v = re;
    ci.push_back((int)I);
    cj.push_back((int)J);
    cv.push_back(v);
    if (I != J && mirror) {
      if (J >= h.M || I >= h.N) return PGS_ERROR_IO;
      T w = v;
      if constexpr (cplx) {
        if (h.storage == 'h') w = std::conj(v);
        else if (h.storage == 'k') w = -v;
      }
      ci.push_back((int)J);
      cj.push_back((int)I);
      cv.push_back(w);
    }
  }
  // COO -> CSR: rows bucketed, COO order kept inside a row
  A = pgs::Csr<T>();
  A.nrows = (int)h.M;
  A.ncols = (int)h.N;
  A.ia.assign(A.nrows + 1, 0);
  for (int i : ci) ++A.ia[i + 1];
  for (int i = 0; i < A.nrows; ++i) A.ia[i + 1] += A.ia[i];
  A.ja.resize(ci.size());
  A.a.resize(ci.size());
  std::vector<int> fill(A.ia.begin(), A.ia.end() - 1);
  for (size_t k = 0; k < ci.size(); ++k) {
    const int d = fill[ci[k]]++;
    A.ja[d] = cj[k];
    A.a[d] = cv[k];
  }
  return 0;
}

template <typename T>
int read_vector(const char* path, int idxin, std::vector<T>& v) {
  constexpr bool cplx = pgs::is_complex<T>::value;
  std::string buf;
  if (!read_file(path, buf)) return PGS_ERROR_IO;
  MmHeader h;
  const long long off = parse_header(buf, h);
  if (off < 0) return PGS_ERROR_IO;
  const bool field_ok = h.field == 'r' || h.field == 'i' || (cplx && h.field == 'c');
  if (!field_ok || !h.coordinate) return PGS_ERROR_IO;
  if (cplx && h.N != 1) return PGS_ERROR_IO;
  v.assign(h.M, T(0.0));
  EntryReader r{buf.c_str() + off, buf.c_str() + buf.size()};
  for (long long k = 0; k < h.nz; ++k) {
    long long I, J;
    double re = 0.0, im = 0.0;
    if (!r.index(I) || !r.index(J) || !r.value(re)) return PGS_ERROR_IO;
    if (h.field == 'c' && !r.value(im)) return PGS_ERROR_IO;
    I -= idxin;
    if (I < 0 || I >= h.M) return PGS_ERROR_IO;
    if constexpr (cplx) v[I] = T(re, im);
    else v[I] = re;
  }
  return 0;
}

template <typename T>
void* copy_out(const std::vector<T>& x, size_t elem) {
  void* p = std::malloc(std::max<size_t>(1, x.size()) * elem);
  if (p && !x.empty()) std::memcpy(p, x.data(), x.size() * elem);
  return p;
}

template <typename T>
int export_matrix(const pgs::Csr<T>& A, int* nrows, int* ncols, int64_t* nnz, int** ia, int** ja, double** a) {
  *ia = (int*)copy_out(A.ia, sizeof(int));
  *ja = (int*)copy_out(A.ja, sizeof(int));
  *a = (double*)copy_out(A.a, sizeof(T));
  if (!*ia || !*ja || !*a) {
    std::free(*ia);
    std::free(*ja);
    std::free(*a);
    *ia = *ja = nullptr;
    *a = nullptr;
    return PGS_ERR_OOM;
  }
  *nrows = A.nrows;
  *ncols = A.ncols;
  *nnz = A.nnz();
  return 0;
}

}  // namespace

extern "C" {

int pgs_read_mm_csr(const char* path, int idxin, int is_complex, int* nrows, int* ncols, int64_t* nnz, int** ia,
                    int** ja, double** a) {
  if (!path || !nrows || !ncols || !nnz || !ia || !ja || !a) return PGS_ERROR_INVALID_PARAM;
  *ia = *ja = nullptr;
  *a = nullptr;
  try {
    if (is_complex) {
      pgs::Csr<pgs::zdouble> A;
      int err = read_matrix(path, idxin, A);
      return err ? err : export_matrix(A, nrows, ncols, nnz, ia, ja, a);
    }
    pgs::Csr<double> A;
    int err = read_matrix(path, idxin, A);
    return err ? err : export_matrix(A, nrows, ncols, nnz, ia, ja, a);
  } catch (const std::bad_alloc&) {
    return PGS_ERR_OOM;
  }
}

int pgs_read_mm_vector(const char* path, int idxin, int is_complex, int* n, double** v) {
  if (!path || !n || !v) return PGS_ERROR_INVALID_PARAM;
  *v = nullptr;
  try {
    int err;
    if (is_complex) {
      std::vector<pgs::zdouble> x;
      err = read_vector(path, idxin, x);
      if (!err) {
        *v = (double*)copy_out(x, sizeof(pgs::zdouble));
        *n = (int)x.size();
      }
    } else {
      std::vector<double> x;
      err = read_vector(path, idxin, x);
      if (!err) {
        *v = (double*)copy_out(x, sizeof(double));
        *n = (int)x.size();
      }
    }
    if (!err && !*v) err = PGS_ERR_OOM;
    return err;
  } catch (const std::bad_alloc&) {
    return PGS_ERR_OOM;
  }
}

}  // extern "C"
