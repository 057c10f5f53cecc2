// Operator files: write the output of the GeMSLR setup to disk and read it
// back, so that a process (a restarted server, a second job on the same
// matrix) skips the setup and goes straight to the device upload.
//
// Layout (little-endian, no padding):
//   "PGSGMSLR"  u32 version (1)  u32 kind (0 double, 1 complex<double>)
//   parameters (each field of GemslrParams in declaration order: int32,
//     double, uint8 for bool, uint64 for the seed)
//   int32 n, int32 nlev, uint8 solve_phase, double setup_seconds[4]
//   csr A, vec<int32> perm, vec<int32> lev_ptr
//   per level: int32 ncomp, vec<int32> dom_ptr,
//     per block: int32 nb, csr L, csr U, vec<T> Dinv,
//     csr E, csr F, csr C, int32 rank, vec<T> W, vec<T> WH, vec<T> Hk
//   uint64 FNV-1a checksum of every byte before it
// with vec<X> = int64 count + count * sizeof(X) bytes and
// csr = int32 nrows, int32 ncols, vec<int32> ia, vec<int32> ja, vec<T> a.
// The level factors Lcat / Ucat / Dcat are rebuilt from the per-block factors
// on load (the same routine the setup runs).
#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/pargemslr_amd.h"
#include "gemslr_host.hpp"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "operator files are little-endian");

namespace pgs {
namespace {

constexpr char kMagic[8] = {'P', 'G', 'S', 'G', 'M', 'S', 'L', 'R'};
constexpr uint32_t kVersion = 1;
constexpr uint64_t kFnvOffset = 1469598103934665603ULL;
constexpr uint64_t kFnvPrime = 1099511628211ULL;

inline uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * kFnvPrime;
  return h;
}

// Calls v(field) for every parameter, in the file's order.
template <class P, class V>
void visit_params(P& p, V& v) {
  v.i32(p.bsolve_top); v.i32(p.bsolve_lower); v.i32(p.global_solve);
  v.i32(p.nlev); v.i32(p.ncomp); v.i32(p.partition); v.i32(p.minsep); v.i32(p.kmin); v.i32(p.kfactor);
  v.flag(p.vertexsep);
  v.f64(p.droptol_B); v.f64(p.droptol_C);
  v.i32(p.rownnz_B); v.i32(p.rownnz_C); v.i32(p.perm_option);
  v.i32(p.lr_option1); v.i32(p.lr_option2); v.i32(p.lr_rank1); v.i32(p.lr_rank2);
  v.f64(p.lr_rank_factor1); v.f64(p.lr_rank_factor2); v.f64(p.lr_arnoldi_factor1); v.f64(p.lr_arnoldi_factor2);
  v.f64(p.lr_tol_eig1); v.f64(p.lr_tol_eig2);
  v.i32(p.lr_maxits1); v.i32(p.lr_maxits2);
  v.f64(p.lr_tr_factor); v.f64(p.lr_tol_orth);
  v.flag(p.lr_rand_init);
  v.i32(p.schur_enable); v.f64(p.schur_tol); v.i32(p.schur_maxits);
  v.u64(p.seed);
}

struct Writer {
  FILE* f;
  uint64_t h = kFnvOffset;
  bool ok = true;
  void bytes(const void* p, size_t n) {
    if (!ok || n == 0) return;
    if (std::fwrite(p, 1, n, f) != n) ok = false;
    h = fnv1a(h, p, n);
  }
  template <class X> void pod(const X& x) { bytes(&x, sizeof(X)); }
  void i32(const int& x) { pod<int32_t>(x); }
  void f64(const double& x) { pod(x); }
  void u64(const uint64_t& x) { pod(x); }
  void flag(const bool& x) { pod<uint8_t>(x ? 1 : 0); }
  template <class X> void vec(const std::vector<X>& v) {
    pod<int64_t>((int64_t)v.size());
    bytes(v.data(), v.size() * sizeof(X));
  }
  template <class T> void csr(const Csr<T>& m) {
    i32(m.nrows);
    i32(m.ncols);
    vec(m.ia);
    vec(m.ja);
    vec(m.a);
  }
};

struct Reader {
  FILE* f;
  int64_t left;  // payload bytes not yet read (the checksum excluded)
  uint64_t h = kFnvOffset;
  bool ok = true;
  void bytes(void* p, size_t n) {
    if (!ok || n == 0) return;
    if ((int64_t)n > left || std::fread(p, 1, n, f) != n) {
      ok = false;
      return;
    }
    left -= (int64_t)n;
    h = fnv1a(h, p, n);
  }
  template <class X> X pod() {
    X x{};
    bytes(&x, sizeof(X));
    return x;
  }
  void i32(int& x) { x = pod<int32_t>(); }
  void f64(double& x) { x = pod<double>(); }
  void u64(uint64_t& x) { x = pod<uint64_t>(); }
  void flag(bool& x) { x = pod<uint8_t>() != 0; }
  // a count is accepted only if the rest of the file can hold it
  template <class X> void vec(std::vector<X>& v) {
    const int64_t n = pod<int64_t>();
    if (!ok || n < 0 || n > left / (int64_t)sizeof(X)) {
      ok = false;
      return;
    }
    v.resize((size_t)n);
    bytes(v.data(), (size_t)n * sizeof(X));
  }
  template <class T> void csr(Csr<T>& m) {
    i32(m.nrows);
    i32(m.ncols);
    vec(m.ia);
    vec(m.ja);
    vec(m.a);
  }
};

template <class T>
bool csr_valid(const Csr<T>& m) {
  if (m.nrows < 0 || m.ncols < 0 || m.ia.size() != (size_t)m.nrows + 1 || m.ia[0] != 0) return false;
  for (int i = 0; i < m.nrows; ++i)
    if (m.ia[i + 1] < m.ia[i]) return false;
  if ((size_t)m.ia[m.nrows] != m.ja.size() || m.a.size() != m.ja.size()) return false;
  for (int j : m.ja)
    if (j < 0 || j >= m.ncols) return false;
  return true;
}

template <class T>
bool csr_shape(const Csr<T>& m, int nrows, int ncols) {
  return csr_valid(m) && m.nrows == nrows && m.ncols == ncols;
}

// strictly lower (lower) or strictly upper triangular pattern
template <class T>
bool strict_triangle(const Csr<T>& m, bool lower) {
  for (int i = 0; i < m.nrows; ++i)
    for (int k = m.ia[i]; k < m.ia[i + 1]; ++k)
      if (lower ? m.ja[k] >= i : m.ja[k] <= i) return false;
  return true;
}

bool offsets_valid(const std::vector<int>& p, int first, int last) {
  if (p.empty() || p.front() != first || p.back() != last) return false;
  for (size_t i = 1; i < p.size(); ++i)
    if (p[i] < p[i - 1]) return false;
  return true;
}

// Structure of a read operator; returns an empty string or what is wrong.
template <class T>
std::string check_op(const HostOp<T>& op) {
  const int n = op.n;
  if (n < 0 || op.nlev < 1) return "sizes";
  if (!csr_shape(op.A, n, n)) return "matrix";
  if (op.perm.size() != (size_t)n) return "permutation";
  {
    std::vector<char> seen(n, 0);
    for (int p : op.perm) {
      if (p < 0 || p >= n || seen[p]) return "permutation";
      seen[p] = 1;
    }
  }
  if (op.lev_ptr.size() != (size_t)op.nlev + 1 || !offsets_valid(op.lev_ptr, 0, n)) return "level offsets";
  const GemslrParams& prm = op.prm;
  if (prm.schur_enable != 0 &&
      (prm.schur_maxits < 1 || prm.schur_maxits > PGS_SCHUR_MAXITS_LIMIT || !(prm.schur_tol >= 0.0)))
    return "inner iteration settings";
  for (int l = 0; l < op.nlev; ++l) {
    const HostLevel<T>& L = op.levels[l];
    const int s = op.lev_ptr[l], e = op.lev_ptr[l + 1], nl = e - s, m = n - e;
    const bool last = l == op.nlev - 1;
    if (L.ncomp < 1 || L.dom_ptr.size() != (size_t)L.ncomp + 1 || !offsets_valid(L.dom_ptr, s, e) ||
        L.ilu.size() != (size_t)L.ncomp)
      return "blocks of level " + std::to_string(l);
    for (int b = 0; b < L.ncomp; ++b) {
      const IluFactor<T>& F = L.ilu[b];
      const int nb = L.dom_ptr[b + 1] - L.dom_ptr[b];
      if (F.n != nb || !csr_shape(F.L, nb, nb) || !csr_shape(F.U, nb, nb) || F.Dinv.size() != (size_t)nb ||
          !strict_triangle(F.L, true) || !strict_triangle(F.U, false))
        return "ILUT factor of level " + std::to_string(l) + " block " + std::to_string(b);
    }
    if (!last) {
      if (!csr_shape(L.E, m, nl) || !csr_shape(L.F, nl, m)) return "couplings of level " + std::to_string(l);
    } else if (!csr_valid(L.E) || !csr_valid(L.F) || L.E.nnz() != 0 || L.F.nnz() != 0) {
      return "couplings of the last level";
    }
    const bool needs_c = l == 0 && prm.schur_enable && !last;
    const bool empty_c = csr_valid(L.C) && L.C.nrows == 0 && L.C.nnz() == 0;
    if (needs_c ? !csr_shape(L.C, m, m) : !(empty_c || csr_shape(L.C, m, m)))
      return "Schur block of level " + std::to_string(l);
    if (L.rank < 0 || (last && L.rank != 0) || (int64_t)L.W.size() != (int64_t)m * L.rank ||
        L.WH.size() != L.W.size() || !(L.Hk.empty() || (int64_t)L.Hk.size() == (int64_t)L.rank * L.rank))
      return "low-rank term of level " + std::to_string(l);
  }
  return std::string();
}

FILE* open_read(const char* path, int64_t& size) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return nullptr;
  if (fseeko(f, 0, SEEK_END) != 0) {
    std::fclose(f);
    return nullptr;
  }
  size = (int64_t)ftello(f);
  if (size < 0 || fseeko(f, 0, SEEK_SET) != 0) {
    std::fclose(f);
    return nullptr;
  }
  return f;
}

}  // namespace

template <typename T>
int op_save(const HostOp<T>& op, const char* path) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return PGS_ERROR_IO;
  Writer w{f};
  w.bytes(kMagic, sizeof(kMagic));
  w.pod<uint32_t>(kVersion);
  w.pod<uint32_t>(is_complex<T>::value ? 1u : 0u);
  GemslrParams prm = op.prm;
  visit_params(prm, w);
  w.i32(op.n);
  w.i32(op.nlev);
  w.flag(op.solve_phase);
  for (double t : op.setup_seconds) w.f64(t);
  w.csr(op.A);
  w.vec(op.perm);
  w.vec(op.lev_ptr);
  for (int l = 0; l < op.nlev; ++l) {
    const HostLevel<T>& L = op.levels[l];
    w.i32(L.ncomp);
    w.vec(L.dom_ptr);
    for (int b = 0; b < L.ncomp; ++b) {
      w.i32(L.ilu[b].n);
      w.csr(L.ilu[b].L);
      w.csr(L.ilu[b].U);
      w.vec(L.ilu[b].Dinv);
    }
    w.csr(L.E);
    w.csr(L.F);
    w.csr(L.C);
    w.i32(L.rank);
    w.vec(L.W);
    w.vec(L.WH);
    w.vec(L.Hk);
  }
  const uint64_t sum = w.h;
  w.pod(sum);
  const bool ok = w.ok && std::fflush(f) == 0;
  if (std::fclose(f) != 0 || !ok) {
    std::remove(tmp.c_str());
    return PGS_ERROR_IO;
  }
  if (std::rename(tmp.c_str(), path) != 0) {
    std::remove(tmp.c_str());
    return PGS_ERROR_IO;
  }
  return 0;
}

int op_file_kind(const char* path, int* kind) {
  int64_t size = 0;
  FILE* f = open_read(path, size);
  if (!f) return PGS_ERROR_IO;
  char magic[8];
  uint32_t hdr[2];
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::fread(hdr, 4, 2, f) == 2;
  std::fclose(f);
  if (!ok || std::string(magic, 8) != std::string(kMagic, 8) || hdr[0] != kVersion || hdr[1] > 1) return PGS_ERROR_IO;
  *kind = (int)hdr[1];
  return 0;
}

template <typename T>
int op_load(const char* path, HostOp<T>& op, std::string* msg) {
  auto fail = [&](const std::string& why) {
    if (msg) *msg = "operator file: " + why;
    return PGS_ERROR_IO;
  };
  int64_t size = 0;
  FILE* f = open_read(path, size);
  if (!f) return fail("cannot open");
  if (size < (int64_t)(sizeof(kMagic) + 8 + 8)) {
    std::fclose(f);
    return fail("too short");
  }
  Reader r{f, size - 8};
  char magic[8];
  r.bytes(magic, 8);
  const uint32_t version = r.pod<uint32_t>();
  const uint32_t kind = r.pod<uint32_t>();
  if (!r.ok || std::string(magic, 8) != std::string(kMagic, 8) || version != kVersion) {
    std::fclose(f);
    return fail("not an operator file of this version");
  }
  if (kind != (is_complex<T>::value ? 1u : 0u)) {
    std::fclose(f);
    return fail("value kind");
  }
  op = HostOp<T>();
  visit_params(op.prm, r);
  r.i32(op.n);
  r.i32(op.nlev);
  r.flag(op.solve_phase);
  for (double& t : op.setup_seconds) r.f64(t);
  r.csr(op.A);
  r.vec(op.perm);
  r.vec(op.lev_ptr);
  // each level takes at least 4 + 8 bytes, each block at least 4 + 3 * 8
  if (!r.ok || op.nlev < 1 || op.nlev > r.left / 12) {
    std::fclose(f);
    return fail("truncated or corrupt header");
  }
  op.levels.resize(op.nlev);
  for (int l = 0; r.ok && l < op.nlev; ++l) {
    HostLevel<T>& L = op.levels[l];
    r.i32(L.ncomp);
    if (!r.ok || L.ncomp < 1 || L.ncomp > r.left / 28) {
      r.ok = false;
      break;
    }
    r.vec(L.dom_ptr);
    L.ilu.resize(L.ncomp);
    for (int b = 0; r.ok && b < L.ncomp; ++b) {
      r.i32(L.ilu[b].n);
      r.csr(L.ilu[b].L);
      r.csr(L.ilu[b].U);
      r.vec(L.ilu[b].Dinv);
    }
    r.csr(L.E);
    r.csr(L.F);
    r.csr(L.C);
    r.i32(L.rank);
    r.vec(L.W);
    r.vec(L.WH);
    r.vec(L.Hk);
  }
  uint64_t stored = 0;
  const bool whole = r.ok && r.left == 0 && std::fread(&stored, 8, 1, f) == 1;
  std::fclose(f);
  if (!whole) return fail("truncated or corrupt");
  if (stored != r.h) return fail("checksum mismatch");
  const std::string bad = check_op(op);
  if (!bad.empty()) return fail("invalid " + bad);
  for (int l = 0; l < op.nlev; ++l) build_level_cat(op, l);
  return 0;
}

template int op_save<double>(const HostOp<double>&, const char*);
template int op_save<zdouble>(const HostOp<zdouble>&, const char*);
template int op_load<double>(const char*, HostOp<double>&, std::string*);
template int op_load<zdouble>(const char*, HostOp<zdouble>&, std::string*);

}  // namespace pgs
