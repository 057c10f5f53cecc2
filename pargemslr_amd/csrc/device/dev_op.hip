// Device-resident GeMSLR operator, workspaces, the preconditioner apply and
// the FGMRES driver, exported through the C ABI of pargemslr_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/pargemslr_amd.h"
#include "../host/gemslr_host.hpp"
#include "../host/trsv_pack.hpp"
#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace pgs;
using namespace pgs::dev;

namespace {

struct HipError {
  hipError_t e;
};

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) throw HipError{_e};          \
  } while (0)

thread_local std::string g_last_hip_error;

inline unsigned nblocks_for(int64_t work, int nt) { return (unsigned)((work + nt - 1) / nt); }

struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  bool own = true;  // false: a view of memory that this object does not free
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() {
    if (p && own) (void)hipFree(p);
  }
  void view(void* q, size_t b) {
    if (p && own) (void)hipFree(p);
    p = q;
    bytes = b;
    own = false;
  }
  void alloc(size_t b) {
    if (p && own) (void)hipFree(p);
    p = nullptr;
    own = true;
    bytes = b;
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) { p = nullptr; throw HipError{e}; }
  }
  template <typename T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

template <typename T>
void upload(DevMem& m, const std::vector<T>& v, int64_t* counter) {
  m.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpy(m.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  if (counter) *counter += (int64_t)(v.size() * sizeof(T));
}

// host value (double or std::complex<double>) -> device value
template <typename To, typename From>
struct HVal {
  static To get(const From& v) { return (To)v; }
};
template <typename R>
struct HVal<cplx<R>, zdouble> {
  static cplx<R> get(const zdouble& v) { return cplx<R>((R)v.real(), (R)v.imag()); }
};
template <typename To, typename From>
To hval(const From& v) { return HVal<To, From>::get(v); }

template <typename To, typename From>
std::vector<To> convert_vec(const std::vector<From>& v) {
  std::vector<To> o(v.size());
  for (size_t i = 0; i < v.size(); ++i) o[i] = hval<To>(v[i]);
  return o;
}

// W (column-major m x k) -> row-major
template <typename To, typename From>
std::vector<To> to_row_major(const std::vector<From>& W, int m, int k) {
  std::vector<To> o((size_t)m * k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < m; ++i) o[(size_t)i * k + j] = hval<To>(W[(size_t)j * m + i]);
  return o;
}

struct DevCsr {
  int nrows = 0, ncols = 0;
  int64_t nnz = 0;
  int lpr = 1;
  int spmm_cols = 4;  // columns per thread of products with 8 or more columns (4 or 8)
  DevMem ia, ja, a;
};

template <typename V, typename T>
void upload_csr(DevCsr& d, const Csr<T>& A, int64_t* counter) {
  d.nrows = A.nrows;
  d.ncols = A.ncols;
  d.nnz = A.nnz();
  upload(d.ia, A.ia, counter);
  upload(d.ja, A.ja, counter);
  upload(d.a, convert_vec<V>(A.a), counter);
  // one-column products: one thread per row (eight entries in flight) up to
  // 12 entries per row on average, else that many lanes per row (128^3
  // Laplacian, 7 entries per row: 41 us against 60 us with 8 lanes per row)
  double avg = A.nrows > 0 ? (double)A.nnz() / A.nrows : 0.0;
  d.lpr = avg <= 12 ? 1 : avg <= 16 ? 16 : avg <= 32 ? 32 : 64;
  // PGS_SPMV_LPR=1, 2, 4, 8 ... fixes the lanes per row of one-column
  // products (1: one thread per row, eight entries in flight)
  if (const char* e = std::getenv("PGS_SPMV_LPR")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32 || v == 64) d.lpr = v;
  }
  // PGS_SPMM_COLS=8: eight columns per thread in products of 8 or more columns
  if (const char* e = std::getenv("PGS_SPMM_COLS")) d.spmm_cols = std::atoi(e) == 8 ? 8 : 4;
}

// Pipelined triangular-solve geometry: 512 threads, 4 lanes per row, 5 slots
// per lane -> steps of at most 128 rows and 20 slots.
constexpr int kTrsvMaxWidth = 20;  // widest factor row the pipelined kernel takes

struct DevTrsv {
  int nblocks = 0, nloc = 0, max_rows = 0, max_steps = 0, max_width = 0;
  int window = 0;  // ring size when the packed columns are ring slots, else 0
  int64_t nnz_padded = 0, nnz_real = 0, nrows_total = 0;
  int64_t max_block_ent = 0, max_block_steps = 0;  // largest entry and step counts of one block
  DevMem row0, step, stepU, ent0, ri0, desc, ent, ri;
  // Half steps for the one-wave ring kernel (64 threads, 4 lanes per row, 16
  // rows per pass) over the same ring packing: a step of more than 16 rows is
  // walked as two consecutive half steps (rows 0-15, then 16-31).  Rows of a
  // step are independent, and a half step reads and writes a subset of its
  // step's positions, so the ring window still holds.
  int64_t max_block_steps_h = 0;
  DevMem step_h, stepU_h, desc_h;
};


template <typename V, typename T>
void upload_trsv(DevTrsv& d, const TrsvPack<T>& P, int64_t* counter) {
  d.nblocks = P.nblocks;
  d.nloc = P.nloc;
  d.max_rows = P.max_block_rows;
  d.max_steps = P.max_steps;
  d.max_width = P.max_width;
  d.window = P.window;
  d.nnz_padded = (int64_t)P.col.size();
  d.nnz_real = P.nnz_real;
  d.nrows_total = (int64_t)P.rows.size();
  d.max_block_ent = d.max_block_steps = 0;
  for (int b = 0; b < P.nblocks; ++b) {
    d.max_block_ent = std::max<int64_t>(d.max_block_ent, P.blk_ent0[b + 1] - P.blk_ent0[b]);
    d.max_block_steps = std::max<int64_t>(d.max_block_steps, P.blk_step[b + 1] - P.blk_step[b]);
  }
  upload(d.row0, P.blk_row0, counter);
  upload(d.step, P.blk_step, counter);
  upload(d.stepU, P.blk_stepU, counter);
  upload(d.ent0, P.blk_ent0, counter);
  upload(d.ri0, P.blk_ri0, counter);
  upload(d.desc, P.desc, counter);
  d.max_block_steps_h = 0;
  int step_rows = 0;
  for (size_t i = 2; i < P.desc.size(); i += 4) step_rows = std::max(step_rows, P.desc[i]);
  if (P.window > 0 && step_rows <= 32) {
    std::vector<int> dh, sh, suh;
    half_steps(P.desc, P.blk_step, P.blk_stepU, dh, sh, suh);
    for (int b = 0; b < P.nblocks; ++b)
      d.max_block_steps_h = std::max<int64_t>(d.max_block_steps_h, sh[b + 1] - sh[b]);
    upload(d.desc_h, dh, counter);
    upload(d.step_h, sh, counter);
    upload(d.stepU_h, suh, counter);
  }
  std::vector<TPair<V>> ent(P.col.size() + 1), ri(P.rows.size() + 1);
  constexpr int64_t kChunk = 1 << 20;
  const int64_t ne = (int64_t)P.col.size(), nr = (int64_t)P.rows.size();
  parallel_for((int)((ne + kChunk - 1) / kChunk), [&](int c) {
    for (int64_t i = c * kChunk, i1 = std::min(ne, (c + 1) * kChunk); i < i1; ++i) {
      ent[i] = TPair<V>();
      ent[i].col = P.window > 0 ? P.col[i] * ring_unit<V>() : P.col[i];  // ring slots as 4-column row offsets
      ent[i].val = hval<V>(P.val[i]);
    }
  });
  parallel_for((int)((nr + kChunk - 1) / kChunk), [&](int c) {
    for (int64_t i = c * kChunk, i1 = std::min(nr, (c + 1) * kChunk); i < i1; ++i) {
      ri[i] = TPair<V>();
      ri[i].col = P.rows[i];
      ri[i].val = hval<V>(P.dinv[i]);
    }
  });
  ent.back() = TPair<V>();
  ri.back() = TPair<V>();
  upload(d.ent, ent, counter);
  upload(d.ri, ri, counter);
}

struct DevLevel {
  int lev_start = 0, nloc = 0, nrem = 0, rank = 0;
  DevTrsv B;
  DevCsr E, F;
  DevCsr C;      // level 0 with the inner Schur iteration: the nrem x nrem lower-right block
  DevMem W, WH;  // row-major nrem x rank
};

template <typename V>
struct DevPrec {
  bool present = false;
  std::vector<std::unique_ptr<DevLevel>> levels;
};

}  // namespace

struct pgs_dev_op_s {
  int device = 0;
  int n = 0;
  int nlev = 0;
  int is_complex = 0;
  int lds_max = 65536;
  int trsv_variant = 0;  // 0: pipelined kernels (ring when packed for it); 1: basic level-set kernel
  // most right-hand-side columns per column group of the ring kernel; 0 =
  // chosen per launch (4, or 8 for 4-byte values once 4 would give four or
  // more workgroups per CU)
  int trsv_cg = 0;
  int trsv_groups = 1;   // column groups per workgroup of the 128-thread ring kernel: 1, 2, or 0 = 2 when
                         // the paired grid fills the device (measured slower at 128^3: kept selectable)
  int num_cus = 256;     // compute units of the device
  int trsv_nt = 128;     // threads of the pipelined kernel
  int trsv_lpr = 4;      // lanes per row (steps hold at most trsv_nt / trsv_lpr rows)
  // ring kernel geometry: 0 = chosen per launch (128x4 or one-wave 64x4 on
  // half steps); 128 / 64 with ring_lpr 4 or 2 force one geometry
  int ring_nt = 0;
  int ring_lpr = 4;
  int lr_mfma = 1;          // low-rank term of FP32 blocks of 16, 32 or 64 columns on the matrix cores
  int usolve = 0;           // 1: the U variant of SolveLevel (no block L solve on the way down)
  int inner_maxits = 0;     // > 0: inner FGMRES on the level-0 Schur complement in every apply
  double inner_tol = 0.0;
  int64_t bytes = 0;
  DevCsr A;  // double or complex<double> values
  DevMem perm;
  std::vector<int> lev_ptr;
  DevPrec<double> p64;
  DevPrec<float> p32;
  DevPrec<cplx<double>> pz64;  // complex operators
  DevPrec<cplx<float>> pz32;
};

struct pgs_ws_s {
  pgs_dev_op_s* op = nullptr;
  int max_nrhs = 0, kdim = 0;
  // apply buffers (per precision; float buffers sized as double for reuse)
  DevMem r0, x0, lrpart, lrt;
  std::vector<std::unique_ptr<DevMem>> zlev;
  // Krylov
  DevMem V, Z, part, hbuf, h2buf, t2, H, cs, sn, rs, coef, upd, st, zflag, nb2, nr2, flag;
  DevMem dc_d, dc_coef;  // delayed CGS2: the step's dot table and update coefficients
  DevMem dc_R;           // delayed CGS2: final basis in terms of the raw slots, (kdim+1)^2 per column
  DevMem dc_vscale;      // delayed CGS2: pending normalization of the newest basis vector, per column
  // inner Schur iteration: basis, preconditioned basis, temporaries (vectors
  // in the apply's precision) and its Hessenberg / Givens state per column
  DevMem in_V, in_Z, in_T, in_U, in_Y, in_h, in_t2, in_H, in_cs, in_sn, in_rs, in_coef, in_upd, in_st, in_flag,
      in_nb2;
  int* flag_host = nullptr;
  int part_blocks = 0;
  bool fused_cgs2 = true;  // CGS2 as three k_basis_reg sweeps instead of four
  int64_t basis_pad = 0;    // values between consecutive Krylov vectors (set at creation)
  int64_t v_bytes = 0;
  // kernel timing by family (pgs_ws_profile): HIP events on the launch
  // stream around every launch, and each launch's algorithmic bytes
  bool prof = false;
  std::vector<hipEvent_t> ev;         // pool, used in start/stop pairs
  std::vector<int> ev_fam;            // family of each recorded pair
  size_t ev_used = 0;
  double fam_ms[PGS_PROFILE_FAMILIES] = {};
  double fam_bytes[PGS_PROFILE_FAMILIES] = {};
  int64_t fam_launches[PGS_PROFILE_FAMILIES] = {};
  // captured inner FGMRES steps (index = step), the settings they were
  // recorded with, and the stream they are recorded on
  std::vector<hipGraphExec_t> graph_exec;
  int graph_key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  hipStream_t cap_stream = nullptr;
  void drop_graphs() {
    for (hipGraphExec_t& g : graph_exec)
      if (g) (void)hipGraphExecDestroy(g);
    graph_exec.clear();
  }
  ~pgs_ws_s() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    drop_graphs();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
  }
};

namespace {

constexpr int kNT = 256;
constexpr int kTrsvNT = 256;
constexpr int kPMAX = 128;  // Krylov dimension + 1 limit of the shared-memory coefficient tables
constexpr int kLrBlocks = 1024;  // most workgroups (partial sums) of the low-rank stage 1
constexpr int kBasisNT = 256;
constexpr int kBasisPM = 64;  // most basis vectors k_basis_reg takes
constexpr int kPartRowsDcgs2 = 2 * 32 + 2;  // outputs per column of one k_dcgs2_dots chunk

template <typename V>
DevPrec<V>& prec_of(pgs_dev_op_s* op);
template <>
DevPrec<double>& prec_of<double>(pgs_dev_op_s* op) { return op->p64; }
template <>
DevPrec<float>& prec_of<float>(pgs_dev_op_s* op) { return op->p32; }
template <>
DevPrec<cplx<double>>& prec_of<cplx<double>>(pgs_dev_op_s* op) { return op->pz64; }
template <>
DevPrec<cplx<float>>& prec_of<cplx<float>>(pgs_dev_op_s* op) { return op->pz32; }

// element size of the operator's vectors (double or complex<double>)
inline size_t vec_es(const pgs_dev_op_s* op) { return op->is_complex ? 16 : 8; }

template <typename VA, typename VX, typename VY, typename VACC>
void launch_spmv(const DevCsr& M, const VX* x, VY* y, const VY* r, VACC alpha, VACC beta, int mode, int nrhs,
                 hipStream_t s) {
  if (M.nrows == 0) return;
  if (nrhs == 1 && M.lpr == 1) {
    const unsigned g = (nblocks_for(M.nrows, kNT) + kXcds - 1) / kXcds * kXcds;
    k_spmvN<VA, VX, VY, VACC><<<g, kNT, 0, s>>>(M.nrows, 1, M.ia.as<int>(), M.ja.as<int>(), M.a.as<VA>(), x, y, r,
                                                alpha, beta, mode);
  } else if (nrhs == 1) {
    const int64_t threads = (int64_t)M.nrows * M.lpr;
    const unsigned g = (nblocks_for(threads, kNT) + kXcds - 1) / kXcds * kXcds;  // see k_spmvN
#define PGS_SPMV_CASE(L) \
  case L: k_spmv1<VA, VX, VY, VACC, L><<<g, kNT, 0, s>>>(M.nrows, M.ia.as<int>(), M.ja.as<int>(), M.a.as<VA>(), x, y, r, alpha, beta, mode); break;
    switch (M.lpr) {
      PGS_SPMV_CASE(2)
      PGS_SPMV_CASE(4)
      PGS_SPMV_CASE(8)
      PGS_SPMV_CASE(16)
      PGS_SPMV_CASE(32)
      default: PGS_SPMV_CASE(64)
    }
#undef PGS_SPMV_CASE
  } else {
    auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
    const bool wide = nrhs % 4 == 0 && al16(x) && al16(y) && (r == nullptr || al16(r));
    const int cw = wide ? (M.spmm_cols == 8 && nrhs % 8 == 0 ? 8 : 4) : 1;
    const int64_t threads = (int64_t)M.nrows * nrhs / cw;
    // a multiple of the XCD count (the SpMV kernels renumber workgroups per
    // XCD; the padding workgroups find no rows)
    const unsigned g = (nblocks_for(threads, kNT) + kXcds - 1) / kXcds * kXcds;
    if (cw == 8)
      k_spmvN4<VA, VX, VY, VACC, 8><<<g, kNT, 0, s>>>(M.nrows, nrhs, M.ia.as<int>(), M.ja.as<int>(), M.a.as<VA>(), x,
                                                      y, r, alpha, beta, mode);
    else if (wide)
      k_spmvN4<VA, VX, VY, VACC><<<g, kNT, 0, s>>>(M.nrows, nrhs, M.ia.as<int>(), M.ja.as<int>(), M.a.as<VA>(), x, y,
                                                   r, alpha, beta, mode);
    else
      k_spmvN<VA, VX, VY, VACC><<<g, kNT, 0, s>>>(M.nrows, nrhs, M.ia.as<int>(), M.ja.as<int>(), M.a.as<VA>(), x, y,
                                                  r, alpha, beta, mode);
  }
  HIPCHK(hipGetLastError());
}

// Folds the recorded event pairs into the family totals (synchronizes them).
void prof_flush(pgs_ws_s* ws) {
  for (size_t i = 0; i + 1 < ws->ev_used; i += 2) {
    HIPCHK(hipEventSynchronize(ws->ev[i + 1]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ws->ev[i], ws->ev[i + 1]));
    ws->fam_ms[ws->ev_fam[i / 2]] += ms;
  }
  ws->ev_used = 0;
  ws->ev_fam.clear();
}

// Start of a timed launch: an event pair is reserved whole, so a flush of a
// full pool never splits a pair.
void prof_begin(pgs_ws_s* ws, hipStream_t s) {
  if (ws->ev_used + 2 > ws->ev.size()) {
    if (ws->ev.size() >= 4096) {
      prof_flush(ws);
    } else {
      for (int k = 0; k < 2; ++k) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        ws->ev.push_back(e);
      }
    }
  }
  HIPCHK(hipEventRecord(ws->ev[ws->ev_used++], s));
}

void prof_end(pgs_ws_s* ws, hipStream_t s, int fam, double bytes) {
  HIPCHK(hipEventRecord(ws->ev[ws->ev_used++], s));
  ws->ev_fam.push_back(fam);
  ws->fam_bytes[fam] += bytes;
  ws->fam_launches[fam] += 1;
}

// A launch (or a few launches of one operation) timed into family `fam` when
// the workspace profiler is on.
template <typename F>
void timed(pgs_ws_s* ws, hipStream_t s, int fam, double bytes, F&& launch) {
  const bool on = ws && ws->prof;
  if (on) prof_begin(ws, s);
  launch();
  if (on) prof_end(ws, s, fam, bytes);
}

// algorithmic bytes of y = M x (+ y or r) over nrhs columns
template <typename VA, typename VX, typename VY>
double spmv_bytes(const DevCsr& M, int nrhs, bool reads_y) {
  return (double)M.nnz * (sizeof(VA) + 4.0) + (M.nrows + 1.0) * 4.0 + (double)M.ncols * nrhs * sizeof(VX) +
         (double)M.nrows * nrhs * sizeof(VY) * (reads_y ? 2.0 : 1.0);
}

template <typename V>
void launch_trsv_impl(pgs_dev_op_s* op, const DevTrsv& T, const V* rhs, V* x, int nrhs, hipStream_t s);

// Level-set triangular solves of one GeMSLR level (family 0 on level 0, 1
// below it when the workspace profiler is on).
template <typename V>
void launch_trsv(pgs_dev_op_s* op, pgs_ws_s* ws, int level, const DevTrsv& T, const V* rhs, V* x, int nrhs,
                 hipStream_t s) {
  const double sv = sizeof(V);
  const double bytes = (double)T.nnz_real * (sv + 4.0) + 2.0 * (T.nloc + 1) * 4.0 + T.nloc * sv +
                       2.0 * (double)T.nloc * nrhs * sv;
  timed(ws, s, level == 0 ? PGS_PROF_TRSV0 : PGS_PROF_TRSV, bytes, [&] { launch_trsv_impl<V>(op, T, rhs, x, nrhs, s); });
}

template <typename V, bool XLDS, int NT, int LPR>
void launch_trsv_pipe_nt(const DevTrsv& T, const V* rhs, V* x, int nrhs, size_t lds, hipStream_t s) {
  // slots per lane so that LPR * EPL covers rows of 20 entries; stage depth
  // keeps the in-flight vector loads (EPL + 2 per stage) under 64
  constexpr int EPL = 20 / LPR;
  constexpr int D = LPR == 4 ? (sizeof(V) == 4 ? 8 : 6) : (sizeof(V) == 4 ? 5 : 4);
  auto kern = k_trsv_pipe<V, XLDS, NT, LPR, EPL, D>;  // LDS limit raised in set_kernel_attrs
  const unsigned grid = (unsigned)T.nblocks * (unsigned)nrhs;
  kern<<<grid, NT, XLDS ? lds : 0, s>>>(T.nblocks, T.row0.as<int>(), T.step.as<int>(), T.stepU.as<int>(),
                                        T.ent0.as<int64_t>(), T.ri0.as<int64_t>(), T.desc.as<int4>(),
                                        T.ent.as<TPair<V>>(), T.ri.as<TPair<V>>(), rhs, x, nrhs);
}

template <typename V, bool XLDS>
void launch_trsv_pipe(int nt, int lpr, const DevTrsv& T, const V* rhs, V* x, int nrhs, size_t lds, hipStream_t s) {
  if (lpr == 2) {
    if (nt == 64) launch_trsv_pipe_nt<V, XLDS, 64, 2>(T, rhs, x, nrhs, lds, s);
    else launch_trsv_pipe_nt<V, XLDS, 128, 2>(T, rhs, x, nrhs, lds, s);
  } else {
    if (nt == 256) launch_trsv_pipe_nt<V, XLDS, 256, 4>(T, rhs, x, nrhs, lds, s);
    else launch_trsv_pipe_nt<V, XLDS, 128, 4>(T, rhs, x, nrhs, lds, s);
  }
}

// Ring kernel geometries: rows of at most kRingWidth entries, LPR lanes per
// row with kRingWidth / LPR slots each; 128 threads and 4 lanes per row (steps
// of at most 32 rows), or one wave of 64 threads with 4 lanes (16 rows) or 2
// lanes (32 rows) per row.
constexpr int kRingWidth = 20;

// Stage depth of the ring kernel's register ring (steps whose entries are in
// flight): the deepest ring whose registers stay within 256 VGPRs, so that
// two waves fit on a SIMD and no stage is moved to AGPRs (a moved stage costs
// register copies that wait for loads in flight).  With 4 lanes per row:
// float 12 (8 for 8-column groups), 8-byte values 6, complex<double> 4; lane
// pairs hold twice the slots per lane: 6 / 4 / 2.
template <typename V, int CG, int LPR = 4>
constexpr int ring_depth() {
  if constexpr (LPR == 2) return sizeof(V) == 4 ? 6 : sizeof(V) == 8 ? 4 : 2;
  return sizeof(V) == 4 ? (CG == 8 ? 8 : 12) : sizeof(V) == 8 ? 6 : 4;
}

template <typename V, int NT, int LPR, int CG, int G = 1>
void launch_trsv_ring_cg(const DevTrsv& T, const V* rhs, V* x, int nrhs, hipStream_t s) {
  constexpr int D = ring_depth<V, CG, LPR>();
  auto kern = k_trsv_ring<V, NT, LPR, kRingWidth / LPR, D, CG, G>;  // LDS limit raised in set_kernel_attrs
  const size_t lds = (size_t)(T.window + 1) * CG * G * sizeof(V);  // a ring + dummy slot per group
  const unsigned grid = (unsigned)T.nblocks * (unsigned)(nrhs / (CG * G));
  // 64 threads with 4 lanes per row take 16 rows per pass: they walk the half steps
  const bool half = NT / LPR < 32;
  const DevMem& step = half ? T.step_h : T.step;
  const DevMem& stepU = half ? T.stepU_h : T.stepU;
  const DevMem& desc = half ? T.desc_h : T.desc;
  kern<<<grid, NT * G, lds, s>>>(T.nblocks, T.window - 1, T.row0.as<int>(), step.as<int>(), stepU.as<int>(),
                                 T.ent0.as<int64_t>(), T.ri0.as<int64_t>(), desc.as<int4>(),
                                 T.ent.as<TPair<V>>(), T.ri.as<TPair<V>>(), rhs, x, nrhs);
}

template <typename V>
void launch_trsv_ring(pgs_dev_op_s* op, const DevTrsv& T, const V* rhs, V* x, int nrhs, hipStream_t s) {
  // the kernel addresses a block's rows of all columns, its entries and its
  // step descriptors with 32-bit byte offsets; lanes with nothing to store
  // aim at offset 0x7ffffff0, which must stay past the block's vector; lanes
  // past a step's rows read up to 32 rows' slabs beyond the step
  if ((int64_t)T.max_rows * nrhs * (int64_t)sizeof(V) > (int64_t)0x7ffffff0 ||
      (T.max_block_ent + 32 * kRingWidth + 1) * (int64_t)sizeof(TPair<V>) >= (int64_t)INT32_MAX ||
      (std::max(T.max_block_steps, T.max_block_steps_h) + 64) * 16 >= (int64_t)INT32_MAX)
    throw std::runtime_error("diagonal block too large for 32-bit offsets");
  if (nrhs <= 0 || (nrhs & (nrhs - 1)) != 0)
    throw std::runtime_error("the ring solve takes a power-of-two number of columns");
  int cg = std::min(nrhs, op->trsv_cg > 0 ? op->trsv_cg : 4);
  // Below one workgroup per CU each group's step chain is the bound, and a
  // narrower group has shorter steps: spread the columns over as many groups
  // as there are CUs (measured at 128^3 level 0, FP32: 4 RHS with 4 columns
  // per group 0.93 ms, with 1 column per group on 256 workgroups ~0.75 ms).
  if (op->trsv_cg == 0)
    while (cg > 1 && (int64_t)T.nblocks * (nrhs / cg) * 2 <= (int64_t)op->num_cus) cg >>= 1;
  // Eight columns per group halve the per-column loads, address arithmetic
  // and reductions but lengthen each step.  They pay when four-column groups
  // would put four or more one-wave workgroups on every CU, and at two per
  // CU when the half steps make a block's chain half again as long as its
  // whole steps (level sets of more than 16 rows), so that 8-column groups
  // of 128 threads on whole steps shorten it.  Measured at 128^3, FP32, ms
  // per launch: level 0 (6,300 half against 3,350 whole steps) at 32 RHS
  // 1.19 with 8-column groups of 128 threads against 1.31 with 4-column
  // one-wave groups, at 64 RHS 1.59 with 8-column one-wave groups against
  // 3.14 with 4 (2.06 with 8 columns and 128 threads), at 16 RHS 0.85 with
  // 4 columns against 1.18 with 8; level 1 (800 half against 700 whole
  // steps) at 32 RHS 0.16 with 4-column one-wave groups against 0.24.
  if (op->trsv_cg == 0 && sizeof(V) == 4 && nrhs >= 8) {
    const int64_t g4 = (int64_t)T.nblocks * (nrhs / 4);
    const bool long_halves = T.desc_h.p && 2 * (int64_t)T.max_block_steps_h >= 3 * (int64_t)T.max_block_steps;
    if (g4 >= 4 * (int64_t)op->num_cus || (g4 >= 2 * (int64_t)op->num_cus && long_halves)) cg = 8;
  }
  if (op->ring_lpr == 2) cg = std::min(cg, 4);  // a lane pair finalizes at most two columns each
  while (cg > 1 && (size_t)(T.window + 1) * cg * sizeof(V) > (size_t)op->lds_max) cg >>= 1;
  // Per launch: at two or more workgroups per CU the CUs' issue and memory
  // pipes bound the solve, and one-wave workgroups on half steps (no barrier
  // per step) do the same work in less time; below that, 128 threads take the
  // steps whole and halve the dependent chain.
  int nt = op->ring_nt;
  if (nt == 0)
    nt = (T.desc_h.p && (int64_t)T.nblocks * (nrhs / cg) >= 2 * (int64_t)op->num_cus) ? 64 : 128;
  if (nt == 64 && op->ring_lpr == 4 && !T.desc_h.p) nt = 128;
  if (nt == 64 && op->ring_lpr == 2) {
    if (cg >= 4) launch_trsv_ring_cg<V, 64, 2, 4>(T, rhs, x, nrhs, s);
    else if (cg >= 2) launch_trsv_ring_cg<V, 64, 2, 2>(T, rhs, x, nrhs, s);
    else launch_trsv_ring_cg<V, 64, 2, 1>(T, rhs, x, nrhs, s);
  } else if (nt == 64) {
    if constexpr (sizeof(V) == 4) {
      if (cg >= 8) {
        launch_trsv_ring_cg<V, 64, 4, 8>(T, rhs, x, nrhs, s);
        return;
      }
    }
    if (cg >= 4) launch_trsv_ring_cg<V, 64, 4, 4>(T, rhs, x, nrhs, s);
    else if (cg == 2) launch_trsv_ring_cg<V, 64, 4, 2>(T, rhs, x, nrhs, s);
    else launch_trsv_ring_cg<V, 64, 4, 1>(T, rhs, x, nrhs, s);
  } else {
    // two 4-column groups per workgroup when both rings fit in LDS and the
    // paired grid still gives every CU a workgroup
    const bool pair = cg >= 4 && op->trsv_groups != 1 && nrhs >= 8 &&
                      (op->trsv_groups == 2 || (int64_t)T.nblocks * (nrhs / 8) >= op->num_cus) &&
                      (size_t)(T.window + 1) * 8 * sizeof(V) <= (size_t)op->lds_max;
    if constexpr (sizeof(V) == 4) {
      if (cg >= 8) {
        launch_trsv_ring_cg<V, 128, 4, 8>(T, rhs, x, nrhs, s);
        return;
      }
    }
    if (pair) launch_trsv_ring_cg<V, 128, 4, 4, 2>(T, rhs, x, nrhs, s);
    else if (cg >= 4) launch_trsv_ring_cg<V, 128, 4, 4>(T, rhs, x, nrhs, s);
    else if (cg == 2) launch_trsv_ring_cg<V, 128, 4, 2>(T, rhs, x, nrhs, s);
    else launch_trsv_ring_cg<V, 128, 4, 1>(T, rhs, x, nrhs, s);
  }
}

template <typename V>
void launch_trsv_impl(pgs_dev_op_s* op, const DevTrsv& T, const V* rhs, V* x, int nrhs, hipStream_t s) {
  if (T.nblocks == 0) return;
  if (T.window > 0) {
    launch_trsv_ring<V>(op, T, rhs, x, nrhs, s);
    HIPCHK(hipGetLastError());
    return;
  }
  const size_t lds = (size_t)T.max_rows * sizeof(V);
  if (!scalar_traits<V>::is_complex && op->trsv_variant == 0 && T.max_width <= kTrsvMaxWidth) {
    if constexpr (!scalar_traits<V>::is_complex) {
      if (lds <= (size_t)op->lds_max) launch_trsv_pipe<V, true>(op->trsv_nt, op->trsv_lpr, T, rhs, x, nrhs, lds, s);
      else launch_trsv_pipe<V, false>(op->trsv_nt, op->trsv_lpr, T, rhs, x, nrhs, lds, s);
    }
  } else {
    k_trsv_direct<V, kTrsvNT><<<T.nblocks, kTrsvNT, 0, s>>>(T.row0.as<int>(), T.step.as<int>(), T.stepU.as<int>(),
                                                            T.ent0.as<int64_t>(), T.ri0.as<int64_t>(),
                                                            T.desc.as<int4>(), T.ent.as<TPair<V>>(),
                                                            T.ri.as<TPair<V>>(), rhs, x, nrhs);
  }
  HIPCHK(hipGetLastError());
}

// Raises the dynamic-LDS limit of every kernel that may need more than 64 KB,
// once per process, so that no launch path calls a HIP function other than
// the launch itself (launches may be captured into graphs).
template <typename K>
void set_lds_limit(K kern, int lds_max) {
  HIPCHK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
}

// The ring kernel addresses its dynamic LDS from byte 0, which holds while it
// has no static LDS.
template <typename K>
void set_ring_attrs(K kern, int lds_max) {
  set_lds_limit(kern, lds_max);
  hipFuncAttributes a;
  HIPCHK(hipFuncGetAttributes(&a, (const void*)kern));
  if (a.sharedSizeBytes != 0) throw std::runtime_error("ring solve kernel with static LDS");
}

template <typename V>
void set_trsv_attrs(int lds_max) {
  constexpr int D1 = ring_depth<V, 1>(), D2 = ring_depth<V, 2>(), D4 = ring_depth<V, 4>();
  set_ring_attrs(k_trsv_ring<V, 128, 4, kRingWidth / 4, D1, 1, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 128, 4, kRingWidth / 4, D2, 2, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 128, 4, kRingWidth / 4, D4, 4, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 128, 4, kRingWidth / 4, D4, 4, 2>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 64, 4, kRingWidth / 4, D1, 1, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 64, 4, kRingWidth / 4, D2, 2, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 64, 4, kRingWidth / 4, D4, 4, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 64, 2, kRingWidth / 2, (ring_depth<V, 1, 2>()), 1, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 64, 2, kRingWidth / 2, (ring_depth<V, 2, 2>()), 2, 1>, lds_max);
  set_ring_attrs(k_trsv_ring<V, 64, 2, kRingWidth / 2, (ring_depth<V, 4, 2>()), 4, 1>, lds_max);
  if constexpr (sizeof(V) == 4) {
    constexpr int D8 = ring_depth<V, 8>();
    set_ring_attrs(k_trsv_ring<V, 128, 4, kRingWidth / 4, D8, 8, 1>, lds_max);
    set_ring_attrs(k_trsv_ring<V, 64, 4, kRingWidth / 4, D8, 8, 1>, lds_max);
  }
  if constexpr (scalar_traits<V>::is_complex) return;  // no pipelined kernel for complex values
  constexpr int DL2 = sizeof(V) == 4 ? 5 : 4;  // the pipelined kernel's stage depths (launch_trsv_pipe_nt)
  constexpr int DP = sizeof(V) == 4 ? 8 : 6;
  set_lds_limit(k_trsv_pipe<V, true, 64, 2, 10, DL2>, lds_max);
  set_lds_limit(k_trsv_pipe<V, true, 128, 2, 10, DL2>, lds_max);
  set_lds_limit(k_trsv_pipe<V, true, 128, 4, 5, DP>, lds_max);
  set_lds_limit(k_trsv_pipe<V, true, 256, 4, 5, DP>, lds_max);
}

template <int PM>
void set_basis_attrs(int lds_max) {
  set_lds_limit(k_basis_reg<kBasisNT, PM, 0, 1>, lds_max);
  set_lds_limit(k_basis_reg<kBasisNT, PM, 1, 1>, lds_max);
  set_lds_limit(k_basis_reg<kBasisNT, PM, 2, 1>, lds_max);
}

void set_kernel_attrs(int lds_max) {
  static std::once_flag once;
  std::call_once(once, [lds_max] {
    set_trsv_attrs<float>(lds_max);
    set_trsv_attrs<double>(lds_max);
    set_trsv_attrs<cplx<float>>(lds_max);
    set_trsv_attrs<cplx<double>>(lds_max);
    set_basis_attrs<8>(lds_max);
    set_basis_attrs<16>(lds_max);
    set_basis_attrs<32>(lds_max);
    set_basis_attrs<64>(lds_max);
    set_lds_limit(k_basis_reg<kBasisNT, 32, 0, 2>, lds_max);
    set_lds_limit(k_basis_reg<kBasisNT, 32, 1, 2>, lds_max);
    // coefficient tables of up to 127 x 64 values in LDS
    set_lds_limit(k_dcgs2_update<double, kBasisNT, 8, 1>, lds_max);
    set_lds_limit(k_dcgs2_update<double, kBasisNT, 16, 1>, lds_max);
    set_lds_limit(k_dcgs2_update<double, kBasisNT, 32, 1>, lds_max);
    set_lds_limit(k_dcgs2_update<double, kBasisNT, 32, 2>, lds_max);
    set_lds_limit(k_dcgs2_update<double, kBasisNT, 64, 2>, lds_max);
    using Z = cplx<double>;
    set_lds_limit(k_dcgs2_update<Z, kBasisNT, 8, 1>, lds_max);
    set_lds_limit(k_dcgs2_update<Z, kBasisNT, 16, 1>, lds_max);
    set_lds_limit(k_dcgs2_update<Z, kBasisNT, 16, 2>, lds_max);
    set_lds_limit(k_dcgs2_update<Z, kBasisNT, 32, 2>, lds_max);
    set_lds_limit(k_dcgs2_dots<Z, kBasisNT, 8, 1>, lds_max);
    set_lds_limit(k_dcgs2_dots<Z, kBasisNT, 8, 2>, lds_max);
    set_lds_limit(k_multi_axpy_norm<double, double, kNT>, lds_max);
    set_lds_limit(k_multi_axpy_norm<cplx<double>, cplx<double>, kNT>, lds_max);
    set_lds_limit(k_update_x<double, double>, lds_max);
    set_lds_limit(k_update_x<float, double>, lds_max);
    set_lds_limit(k_update_x<cplx<double>, cplx<double>>, lds_max);
    set_lds_limit(k_update_x<cplx<float>, cplx<double>>, lds_max);
  });
}

// The low-rank term of a block of 16, 32 or 64 FP32 columns on the matrix
// cores (k_lr_dot_mfma, k_lr_final, k_lr_update_mfma); k <= 32.
template <int KT, int CT>
void launch_lowrank_mfma_kt(const DevLevel& L, float* zl, float* part, float* t, int nrhs, hipStream_t s) {
  const int m = L.nrem, k = L.rank;
  // stage 1: at most kLrBlocks workgroups of four waves, rows per wave a
  // multiple of 4 and at least 64
  int64_t rpw = std::max<int64_t>(64, ((int64_t)m + 4LL * kLrBlocks - 1) / (4LL * kLrBlocks));
  rpw = (rpw + 3) / 4 * 4;
  const int nblk = (int)(((int64_t)m + 4 * rpw - 1) / (4 * rpw));
  k_lr_dot_mfma<KT, CT><<<nblk, 256, 0, s>>>(m, k, nrhs, L.W.as<float>(), zl, part, (int)rpw);
  k_lr_final<float><<<(k * nrhs + 3) / 4, 256, 0, s>>>(nblk, k, nrhs, part, t);
  const int64_t tiles = ((int64_t)m + 15) / 16;
  k_lr_update_mfma<KT, CT><<<(unsigned)((tiles + 3) / 4), 256, 0, s>>>(m, k, nrhs, L.WH.as<float>(), t, zl);
  HIPCHK(hipGetLastError());
}

void launch_lowrank_mfma(pgs_ws_s* ws, const DevLevel& L, float* zl, int nrhs, hipStream_t s) {
  float* part = ws->lrpart.as<float>();
  float* t = ws->lrt.as<float>();
  const bool k16 = L.rank <= 16;
  switch (nrhs) {
    case 16: k16 ? launch_lowrank_mfma_kt<1, 1>(L, zl, part, t, nrhs, s) : launch_lowrank_mfma_kt<2, 1>(L, zl, part, t, nrhs, s); break;
    case 32: k16 ? launch_lowrank_mfma_kt<1, 2>(L, zl, part, t, nrhs, s) : launch_lowrank_mfma_kt<2, 2>(L, zl, part, t, nrhs, s); break;
    default: k16 ? launch_lowrank_mfma_kt<1, 4>(L, zl, part, t, nrhs, s) : launch_lowrank_mfma_kt<2, 4>(L, zl, part, t, nrhs, s); break;
  }
}

template <typename V>
void launch_lowrank(pgs_ws_s* ws, const DevLevel& L, V* zl, int nrhs, hipStream_t s) {
  const int m = L.nrem, k = L.rank;
  if (k <= 0 || m <= 0) return;
  if constexpr (std::is_same<V, float>::value) {
    if (ws->op->lr_mfma && k <= 32 && (nrhs == 16 || nrhs == 32 || nrhs == 64)) {
      launch_lowrank_mfma(ws, L, zl, nrhs, s);
      return;
    }
  }
  V* part = ws->lrpart.as<V>();
  V* t = ws->lrt.as<V>();
  int nblk;
  if (k <= 32) {
    // register partial sums, up to kLrBlocks workgroups
    nblk = (int)std::min<int64_t>(kLrBlocks, ((int64_t)m * nrhs + kNT - 1) / kNT);
    const size_t lds = sizeof(V) * (size_t)(kNT / 64) * k * nrhs;
    if (k <= 8) k_lr_dot_reg<V, kNT, 8><<<nblk, kNT, lds, s>>>(m, k, nrhs, L.W.as<V>(), zl, part);
    else if (k <= 16) k_lr_dot_reg<V, kNT, 16><<<nblk, kNT, lds, s>>>(m, k, nrhs, L.W.as<V>(), zl, part);
    else k_lr_dot_reg<V, kNT, 32><<<nblk, kNT, lds, s>>>(m, k, nrhs, L.W.as<V>(), zl, part);
  } else {
    const int rows_per_blk = std::max(4096, (m + kLrBlocks - 1) / kLrBlocks);
    nblk = (m + rows_per_blk - 1) / rows_per_blk;
    k_lr_dot<V, kNT><<<nblk, kNT, 0, s>>>(m, k, nrhs, L.W.as<V>(), zl, part, rows_per_blk);
  }
  k_lr_final<V><<<(k * nrhs + 3) / 4, 256, 0, s>>>(nblk, k, nrhs, part, t);
  k_lr_update<V><<<nblocks_for((int64_t)m * nrhs, kNT), kNT, 0, s>>>(m, k, nrhs, L.WH.as<V>(), t, zl);
  HIPCHK(hipGetLastError());
}

// out[g] (+)= sum of nblk partials of p*nrhs values of S (a complex value is
// summed as its two doubles)
template <typename S>
void dot_final(int nblk, int p, int nrhs, const S* part, S* out, int accumulate, hipStream_t s) {
  const int pp = p * (int)(sizeof(S) / sizeof(double));
  k_dot_final_wave<<<(pp * nrhs + 3) / 4, 256, 0, s>>>(nblk, pp, nrhs, reinterpret_cast<const double*>(part),
                                                       reinterpret_cast<double*>(out), accumulate);
}

// multi-dot over the first p vectors of V against w: out[j*nrhs+c] = V_j^H w
template <typename VV, typename VW, typename S>
void multidot(pgs_ws_s* ws, int64_t n, int nrhs, int p, const VV* V, int64_t ldv, const VW* w, S* out,
              int accumulate, hipStream_t s) {
  const int64_t rows_per_blk = (n + ws->part_blocks - 1) / ws->part_blocks;
  const int nblk = (int)((n + rows_per_blk - 1) / rows_per_blk);
  k_multidot<VV, VW, S, kNT><<<nblk, kNT, 0, s>>>(n, nrhs, p, V, ldv, w, ws->part.as<S>(), rows_per_blk);
  dot_final<S>(nblk, p, nrhs, ws->part.as<S>(), out, accumulate, s);
  HIPCHK(hipGetLastError());
}

// out[c] = ||x(:, c)||^2
template <typename S>
void colnorm2(pgs_ws_s* ws, int64_t n, int nrhs, const S* x, double* out, hipStream_t s) {
  const int64_t rows_per_blk = (n + ws->part_blocks - 1) / ws->part_blocks;
  const int nblk = (int)((n + rows_per_blk - 1) / rows_per_blk);
  k_colnorm2<S, kNT><<<nblk, kNT, 0, s>>>(n, nrhs, x, ws->part.as<double>(), rows_per_blk);
  k_dot_final_wave<<<(nrhs + 3) / 4, 256, 0, s>>>(nblk, 1, nrhs, ws->part.as<double>(), out, 0);
  HIPCHK(hipGetLastError());
}

// w -= V(:, 0:p) h; with t2, t2[c] = ||w(:, c)||^2 of the result.  V, w in
// VV, h and the arithmetic in S.
template <typename VV, typename S = VV>
void multi_axpy(pgs_ws_s* ws, int64_t n, int nrhs, int p, const VV* V, int64_t ldv, const S* h, VV* w, double* t2,
                hipStream_t s) {
  // enough workgroups to keep every CU busy on the streaming update
  const int nblk = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n * nrhs + kNT - 1) / kNT));
  const size_t shm = sizeof(S) * (size_t)p * nrhs + sizeof(double) * (kNT / 64) * nrhs;
  k_multi_axpy_norm<VV, S, kNT><<<nblk, kNT, shm, s>>>(n, nrhs, p, V, ldv, h, w,
                                                       t2 ? ws->part.as<double>() : nullptr);
  if (t2) k_dot_final_wave<<<(nrhs + 3) / 4, 256, 0, s>>>(nblk, 1, nrhs, ws->part.as<double>(), t2, 0);
  HIPCHK(hipGetLastError());
}

template <typename V>
void solve_level(pgs_dev_op_s* op, pgs_ws_s* ws, DevPrec<V>& P, int l, const V* rhs, V* x, int nrhs, hipStream_t s);

// scalar of the inner iteration's dots, Hessenberg and rotations: double for
// real vectors (float or double), complex<double> for complex vectors
template <typename V>
struct InnerScalar {
  using type = double;
};
template <typename R>
struct InnerScalar<cplx<R>> {
  using type = cplx<double>;
};

// x = inner FGMRES on the level-0 Schur complement, x0 = 0 (the solve phase of
// SolveLevel with inner iterations, gemslr.cpp:3556-3563; FlexGmresClass::Solve,
// fgmres.hpp:263-558, with Krylov dimension = iteration limit = K and relative
// tolerance inner_tol; PargemslrMgs).  Operator: S z = C z - E B^{-1} F z
// (GemslrClass::SchurMatVec, gemslr.cpp:3799-3955); preconditioner:
// SolveLevel(1, v + W H W^H v) (GemslrSchurMatrixClass::Solve, gemslr.cpp:292-320).
// All K steps are launched for every column; a column that has converged (or
// broken down) is masked by its state as in the batched outer FGMRES, so the
// result per column is that of the reference's early exit, and no launch
// waits for the host (the steps may be captured into graphs).
template <typename V>
void inner_schur(pgs_dev_op_s* op, pgs_ws_s* ws, DevPrec<V>& P, const V* b, V* x, int nrhs, hipStream_t s) {
  using S = typename InnerScalar<V>::type;
  const DevLevel& L0 = *P.levels[0];
  const int64_t m = L0.nrem;
  const int K = op->inner_maxits;
  const int64_t ldv = m * nrhs;
  const unsigned nb = nblocks_for(ldv, kNT);
  V* Vb = ws->in_V.as<V>();
  V* Zb = ws->in_Z.as<V>();
  V* T = ws->in_T.as<V>();
  V* U = ws->in_U.as<V>();
  V* Y = ws->in_Y.as<V>();
  S* h = ws->in_h.as<S>();
  double* t2 = ws->in_t2.as<double>();
  S* H = ws->in_H.as<S>();
  S* rs = ws->in_rs.as<S>();
  KState* st = ws->in_st.as<KState>();
  // ||b||; x0 = 0, so the residual is b and normr = normb
  colnorm2<V>(ws, m, nrhs, b, ws->in_nb2.as<double>(), s);
  k_fgmres_init<<<1, 64, 0, s>>>(nrhs, ws->in_nb2.as<double>(), ws->in_nb2.as<double>(), op->inner_tol, 0, st,
                                 ws->in_flag.as<double>());
  k_fill_zero<V><<<nb, kNT, 0, s>>>(ldv, x);
  // v0 = b / normb
  k_copy_conv<V, V><<<nb, kNT, 0, s>>>(ldv, b, Vb);
  k_cycle_start<S><<<1, 64, 0, s>>>(nrhs, K, st, rs);
  k_scale_cols<V><<<nb, kNT, 0, s>>>(m, nrhs, st, Vb);
  HIPCHK(hipGetLastError());
  for (int i = 1; i <= K; ++i) {
    V* vprev = Vb + (int64_t)(i - 1) * ldv;
    V* w = Vb + (int64_t)i * ldv;
    V* zc = Zb + (int64_t)(i - 1) * ldv;
    // z = SolveLevel(1, v + W H W^H v)
    k_copy_conv<V, V><<<nb, kNT, 0, s>>>(ldv, vprev, T);
    launch_lowrank<V>(ws, L0, T, nrhs, s);
    solve_level<V>(op, ws, P, 1, T, zc, nrhs, s);
    // w = -E B^{-1} F z, then w += C z (the order of SchurMatVec)
    launch_spmv<V, V, V, V>(L0.F, zc, U, (const V*)nullptr, V(1.0), V(0.0), 0, nrhs, s);
    launch_trsv<V>(op, ws, 0, L0.B, U, Y, nrhs, s);
    launch_spmv<V, V, V, V>(L0.E, Y, w, (const V*)nullptr, V(-1.0), V(0.0), 0, nrhs, s);
    launch_spmv<V, V, V, V>(L0.C, zc, w, (const V*)nullptr, V(1.0), V(1.0), 0, nrhs, s);
    // modified Gram-Schmidt against v_0..v_{i-1}; ||w||^2 with the last update
    for (int j = 0; j < i; ++j) {
      multidot<V, V, S>(ws, m, nrhs, 1, Vb + (int64_t)j * ldv, ldv, w, h + (int64_t)j * nrhs, 0, s);
      multi_axpy<V, S>(ws, m, nrhs, 1, Vb + (int64_t)j * ldv, ldv, h + (int64_t)j * nrhs, w,
                       j == i - 1 ? t2 : nullptr, s);
    }
    k_givens<S><<<1, 64, 0, s>>>(nrhs, K, i, K, h, t2, H, ws->in_cs.as<S>(), ws->in_sn.as<S>(), rs, st);
    k_scale_cols<V><<<nb, kNT, 0, s>>>(m, nrhs, st, w);
    HIPCHK(hipGetLastError());
  }
  // x = Z y over each column's steps
  k_backsolve<S><<<1, 64, 0, s>>>(nrhs, K, H, rs, st, ws->in_coef.as<S>(), ws->in_upd.as<int>());
  k_update_x<V, S, V><<<nb, kNT, sizeof(S) * K * nrhs, s>>>(m, nrhs, K, Zb, ldv, ws->in_coef.as<S>(),
                                                            ws->in_upd.as<int>(), x);
  HIPCHK(hipGetLastError());
}

// SolveLevel (GemslrClass::SolveLevel, gemslr.cpp:3474-3625) on the device.
template <typename V>
void solve_level(pgs_dev_op_s* op, pgs_ws_s* ws, DevPrec<V>& P, int l, const V* rhs, V* x, int nrhs, hipStream_t s) {
  DevLevel& L = *P.levels[l];
  if (l == op->nlev - 1) {
    launch_trsv<V>(op, ws, l, L.B, rhs, x, nrhs, s);
    return;
  }
  const int nl = L.nloc;
  V* z = ws->zlev[l]->as<V>();
  V* zl = z + (int64_t)nl * nrhs;
  if (op->usolve) {
    // the U variant (kGemslrUSolve, gemslr.cpp:3527-3532): z_l = rhs_l
    const int64_t cnt = (int64_t)L.nrem * nrhs;
    k_copy_conv<V, V><<<nblocks_for(cnt, kNT), kNT, 0, s>>>(cnt, rhs + (int64_t)nl * nrhs, zl);
    HIPCHK(hipGetLastError());
  } else {
    // z_u = B^{-1} rhs_u
    launch_trsv<V>(op, ws, l, L.B, rhs, z, nrhs, s);
    // z_l = rhs_l - E z_u
    timed(ws, s, PGS_PROF_SPMV_EF, spmv_bytes<V, V, V>(L.E, nrhs, true), [&] {
      launch_spmv<V, V, V, V>(L.E, z, zl, rhs + (int64_t)nl * nrhs, (V)1, (V)0, 1, nrhs, s);
    });
  }
  if (l == 0 && op->inner_maxits > 0) {
    // x_l = inner FGMRES on the Schur complement (gemslr.cpp:3556-3563)
    inner_schur<V>(op, ws, P, zl, x + (int64_t)nl * nrhs, nrhs, s);
  } else {
    // z_l += W H W^H z_l
    if (L.rank > 0 && L.nrem > 0)
      timed(ws, s, PGS_PROF_LOWRANK,
            (2.0 * L.rank + 3.0 * nrhs) * (double)L.nrem * sizeof(V), [&] { launch_lowrank<V>(ws, L, zl, nrhs, s); });
    // x_l = SolveLevel(l+1, z_l)
    solve_level<V>(op, ws, P, l + 1, zl, x + (int64_t)nl * nrhs, nrhs, s);
  }
  // z_u = rhs_u - F x_l
  timed(ws, s, PGS_PROF_SPMV_EF, spmv_bytes<V, V, V>(L.F, nrhs, true), [&] {
    launch_spmv<V, V, V, V>(L.F, x + (int64_t)nl * nrhs, z, rhs, (V)1, (V)0, 1, nrhs, s);
  });
  // x_u = B^{-1} z_u
  launch_trsv<V>(op, ws, l, L.B, z, x, nrhs, s);
}

// z = M^{-1} v with v in VI (the operator's double or complex<double>), z in
// VZ, apply arithmetic V; cscale (may be null) scales v's columns first.
template <typename V, typename VI, typename VZ>
void apply_impl(pgs_dev_op_s* op, pgs_ws_s* ws, const VI* v, VZ* z, int nrhs, hipStream_t s,
                const double* cscale = nullptr) {
  DevPrec<V>& P = prec_of<V>(op);
  if (!P.present) throw std::runtime_error("operator was uploaded without this precision");
  const int64_t n = op->n;
  V* r0 = ws->r0.as<V>();
  V* x0 = ws->x0.as<V>();
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool wide = nrhs % 4 == 0 && al16(v) && al16(z);  // workspace buffers are allocation-aligned
  timed(ws, s, PGS_PROF_PERM, (double)n * nrhs * (sizeof(VI) + sizeof(V)) + 4.0 * n, [&] {
    if (wide) k_gather4<VI, V><<<nblocks_for(n * nrhs / 4, kNT), kNT, 0, s>>>(n, nrhs, op->perm.as<int>(), v, r0, cscale);
    else k_gather<VI, V><<<nblocks_for(n * nrhs, kNT), kNT, 0, s>>>(n, nrhs, op->perm.as<int>(), v, r0, cscale);
  });
  HIPCHK(hipGetLastError());
  solve_level<V>(op, ws, P, 0, r0, x0, nrhs, s);
  timed(ws, s, PGS_PROF_PERM, (double)n * nrhs * (sizeof(V) + sizeof(VZ)) + 4.0 * n, [&] {
    if (wide) k_scatter4<V, VZ><<<nblocks_for(n * nrhs / 4, kNT), kNT, 0, s>>>(n, nrhs, op->perm.as<int>(), x0, z);
    else k_scatter<V, VZ><<<nblocks_for(n * nrhs, kNT), kNT, 0, s>>>(n, nrhs, op->perm.as<int>(), x0, z);
  });
  HIPCHK(hipGetLastError());
}

// Uploads A, the permutation and every level (factors packed for the level
// solves, E, F, W, WH) in the precisions of prec_mask: V64 for PGS_PREC_FP64,
// V32 for PGS_PREC_FP32.
template <typename V64, typename V32, typename T>
void upload_op(pgs_dev_op_s* op, const HostOp<T>& H, int prec_mask, bool all_c = false) {
  op->n = H.n;
  op->nlev = H.nlev;
  op->lev_ptr = H.lev_ptr;
  if (H.prm.schur_enable && H.nlev >= 2) {
    op->inner_maxits = H.prm.schur_maxits;
    op->inner_tol = H.prm.schur_tol;
  }
  op->usolve = H.prm.global_solve == 1 ? 1 : 0;
  upload_csr<V64>(op->A, H.A, &op->bytes);
  upload(op->perm, H.perm, &op->bytes);
  // ring packing (ring kernel) unless another kernel geometry was asked for;
  // the ring holds at least one column of the operator's values
  const bool ring = op->trsv_variant == 0 && op->trsv_nt == 128 && op->trsv_lpr == 4;
  // ring steps hold at most 32 rows: 128 threads x 4 lanes or 64 x 2 take a
  // step whole, 64 x 4 takes it in two half steps
  const int max_rows = ring ? 32 : op->trsv_nt / op->trsv_lpr;
  int max_window = 0;
  if (ring) {
    max_window = 1;
    while ((2 * max_window + 1) * (int)sizeof(V64) <= op->lds_max) max_window *= 2;  // + dummy slot
  }
  std::vector<TrsvPack<T>> packs;
  for (int l = 0; l < H.nlev; ++l)
    packs.push_back(trsv_pack(H.levels[l], H.lev_ptr[l], H.nloc(l), max_rows, max_window, kRingWidth));
  auto build = [&](auto& P, auto tag) {
    using V = decltype(tag);
    P.present = true;
    P.levels.clear();
    for (int l = 0; l < H.nlev; ++l) {
      const HostLevel<T>& HL = H.levels[l];
      std::unique_ptr<DevLevel> L(new DevLevel());
      L->lev_start = H.lev_ptr[l];
      L->nloc = H.nloc(l);
      L->nrem = H.nrem(l);
      upload_trsv<V>(L->B, packs[l], &op->bytes);
      if (l + 1 < H.nlev) {
        upload_csr<V>(L->E, HL.E, &op->bytes);
        upload_csr<V>(L->F, HL.F, &op->bytes);
        L->rank = HL.rank;
        if (HL.rank > 0) {
          upload(L->W, to_row_major<V>(HL.W, L->nrem, HL.rank), &op->bytes);
          upload(L->WH, to_row_major<V>(HL.WH, L->nrem, HL.rank), &op->bytes);
        }
        // C of level 0 for the inner Schur iteration; every level's C for the
        // device low-rank build of a setup
        if ((l == 0 && op->inner_maxits > 0) || (all_c && HL.C.nrows > 0)) upload_csr<V>(L->C, HL.C, &op->bytes);
      }
      P.levels.push_back(std::move(L));
    }
  };
  if (prec_mask & PGS_PREC_FP64) build(prec_of<V64>(op), V64(0));
  if (prec_mask & PGS_PREC_FP32) build(prec_of<V32>(op), V32(0));
}

// Krylov vectors (and their preconditioned images) lie kBasisPad values
// apart beyond their length: with a stride of exactly n * nrhs values (a
// power of two times n) the dozens of vectors one sweep streams start on the
// same memory channels.  At 2^21 rows x 32 columns the delayed-CGS2 sweeps
// ran 2-7 % faster with the padded stride (the update over 40 vectors 4.50
// against 4.20 TB/s, the dots over 24 vectors 6.06 against 5.95 TB/s).
constexpr int64_t kBasisPad = 320;

// Byte sizes of every buffer of a workspace (nr columns, a power of two; Krylov
// dimension kdim): one list serves the allocation and pgs_ws_estimate.
struct WsSizes {
  size_t vb = 0;  // one vector block of the operator's values
  int part_blocks = 0;
  std::vector<size_t> zlev;
  std::vector<std::pair<DevMem pgs_ws_s::*, size_t>> bufs;
  size_t total() const {
    size_t t = 0;
    for (size_t b : zlev) t += b;
    for (const auto& e : bufs) t += e.second;
    return t;
  }
};

// lowrank_build: only what the device low-rank build's level solves,
// multi-dots and norms use (level temporaries, partial sums, the low-rank
// term's partials); no Krylov bases, Hessenberg state or inner-iteration
// buffers.
WsSizes ws_sizes(const pgs_dev_op_s* op, int nr, int kdim, bool lowrank_build = false) {
  WsSizes z;
  const int64_t n = op->n;
  const size_t es = vec_es(op);  // bytes of one value of the operator (double or complex<double>)
  const size_t vb = (size_t)n * nr * es;
  z.vb = vb;
  for (int l = 0; l < op->nlev; ++l) z.zlev.push_back((size_t)(n - op->lev_ptr[l]) * nr * es);
  int maxrank = 0;
  for (int l = 0; l + 1 < op->nlev; ++l) {
    const auto& lv = op->p64.present ? op->p64.levels : op->p32.present ? op->p32.levels
                     : op->pz64.present ? op->pz64.levels : op->pz32.levels;
    maxrank = std::max(maxrank, lv[l]->rank);
  }
  // dot products use at most 1024 workgroups (four per CU); the partials
  // are summed by one wave per output
  const int64_t rows_per_blk = std::max<int64_t>(256, (n + 1023) / 1024);
  z.part_blocks = (int)std::max<int64_t>(1, (n + rows_per_blk - 1) / rows_per_blk);
  const size_t ik = (size_t)kdim;
  auto add = [&](DevMem pgs_ws_s::*m, size_t b) { z.bufs.push_back({m, b}); };
  if (lowrank_build) {
    add(&pgs_ws_s::lrpart, (size_t)std::max(1, (kLrBlocks + 1) * maxrank * nr) * es);
    add(&pgs_ws_s::lrt, (size_t)std::max(1, maxrank * nr) * es);
    add(&pgs_ws_s::part, (size_t)z.part_blocks * std::max(kdim + 1, kPartRowsDcgs2) * nr * es);
    add(&pgs_ws_s::flag, 16);
    return z;
  }
  add(&pgs_ws_s::r0, vb);
  add(&pgs_ws_s::x0, vb);
  add(&pgs_ws_s::lrpart, (size_t)std::max(1, (kLrBlocks + 1) * maxrank * nr) * es);
  add(&pgs_ws_s::lrt, (size_t)std::max(1, maxrank * nr) * es);
  add(&pgs_ws_s::V, (ik + 1) * (vb + kBasisPad * es));
  add(&pgs_ws_s::Z, ik * (vb + kBasisPad * es));
  // partial rows per workgroup: kdim + 1 dots, or 2 * 32 + 2 of a delayed-CGS2 dot chunk
  add(&pgs_ws_s::part, (size_t)z.part_blocks * std::max(kdim + 1, kPartRowsDcgs2) * nr * es);
  add(&pgs_ws_s::hbuf, (ik + 1) * nr * es);
  add(&pgs_ws_s::h2buf, (ik + 1) * nr * es);
  add(&pgs_ws_s::t2, (size_t)nr * sizeof(double));
  add(&pgs_ws_s::H, (size_t)nr * (ik + 1) * ik * es);
  add(&pgs_ws_s::cs, (size_t)nr * ik * es);
  add(&pgs_ws_s::sn, (size_t)nr * ik * es);
  add(&pgs_ws_s::rs, (size_t)nr * (ik + 1) * es);
  add(&pgs_ws_s::coef, (ik + 1) * nr * es);
  add(&pgs_ws_s::upd, (size_t)nr * sizeof(int));
  add(&pgs_ws_s::dc_d, (2 * ik + 4) * nr * es);
  add(&pgs_ws_s::dc_coef, (2 * ik + 4) * nr * es);
  add(&pgs_ws_s::dc_R, (ik + 1) * (ik + 1) * nr * es);
  add(&pgs_ws_s::dc_vscale, (size_t)nr * sizeof(double));
  add(&pgs_ws_s::st, (size_t)nr * sizeof(KState));
  add(&pgs_ws_s::zflag, (size_t)nr * sizeof(double));
  add(&pgs_ws_s::nb2, (size_t)nr * sizeof(double));
  add(&pgs_ws_s::nr2, (size_t)nr * sizeof(double));
  add(&pgs_ws_s::flag, 16);
  if (op->inner_maxits > 0) {
    // inner Schur iteration: basis, preconditioned basis and temporaries of
    // the level-0 remainder (m0) and of level 0 (nl0); its scalars are
    // double or complex<double>, as the operator's values
    const size_t K = (size_t)op->inner_maxits;
    const size_t m0 = (size_t)(n - op->lev_ptr[1]), nl0 = (size_t)std::max(1, op->lev_ptr[1]);
    add(&pgs_ws_s::in_V, (K + 1) * m0 * nr * es);
    add(&pgs_ws_s::in_Z, K * m0 * nr * es);
    add(&pgs_ws_s::in_T, m0 * nr * es);
    add(&pgs_ws_s::in_U, nl0 * nr * es);
    add(&pgs_ws_s::in_Y, nl0 * nr * es);
    add(&pgs_ws_s::in_h, K * nr * es);
    add(&pgs_ws_s::in_t2, (size_t)nr * sizeof(double));
    add(&pgs_ws_s::in_H, (size_t)nr * (K + 1) * K * es);
    add(&pgs_ws_s::in_cs, (size_t)nr * K * es);
    add(&pgs_ws_s::in_sn, (size_t)nr * K * es);
    add(&pgs_ws_s::in_rs, (size_t)nr * (K + 1) * es);
    add(&pgs_ws_s::in_coef, K * nr * es);
    add(&pgs_ws_s::in_upd, (size_t)nr * sizeof(int));
    add(&pgs_ws_s::in_st, (size_t)nr * sizeof(KState));
    add(&pgs_ws_s::in_flag, (size_t)nr * sizeof(double));
    add(&pgs_ws_s::in_nb2, (size_t)nr * sizeof(double));
  }
  return z;
}

int set_err(const HipError& e) {
  g_last_hip_error = hipGetErrorString(e.e);
  if (e.e == hipErrorOutOfMemory) return PGS_ERR_OOM;
  return PGS_ERR_HIP;
}

bool pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

}  // namespace

extern "C" {

int pgs_device_count(int* count) {
  if (!count) return PGS_ERROR_INVALID_PARAM;
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) { *count = 0; return PGS_ERR_NO_DEVICE; }
  *count = c;
  return 0;
}

int pgs_dev_op_create(pgs_gemslr_t hop, int device, int prec_mask, pgs_dev_op_t* out) {
  if (!hop || !out) return PGS_ERROR_INVALID_PARAM;
  *out = nullptr;
  pgs_gemslr_s* h = reinterpret_cast<pgs_gemslr_s*>(hop);
  if ((prec_mask & (PGS_PREC_FP64 | PGS_PREC_FP32)) == 0) return PGS_ERROR_INVALID_PARAM;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PGS_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return PGS_ERROR_INVALID_PARAM;
  std::unique_ptr<pgs_dev_op_s> op(new pgs_dev_op_s());
  try {
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    op->device = device;
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    op->lds_max = lds;
    HIPCHK(hipDeviceGetAttribute(&op->num_cus, hipDeviceAttributeMultiprocessorCount, device));
    set_kernel_attrs(lds);
    if (const char* v = std::getenv("PGS_TRSV_VARIANT")) op->trsv_variant = std::atoi(v);
    if (const char* v = std::getenv("PGS_LR_MFMA")) op->lr_mfma = std::atoi(v) != 0;
    if (const char* v = std::getenv("PGS_TRSV_NT")) {
      const int nt = std::atoi(v);
      if (nt == 64 || nt == 128 || nt == 256) op->trsv_nt = nt;
    }
    if (const char* v = std::getenv("PGS_TRSV_LPR")) {
      const int lpr = std::atoi(v);
      if (lpr == 2 || lpr == 4) op->trsv_lpr = lpr;
    }
    if (op->trsv_lpr == 2 && op->trsv_nt == 256) op->trsv_nt = 128;
    if (op->trsv_lpr == 4 && op->trsv_nt == 64) op->trsv_nt = 128;
    if (const char* v = std::getenv("PGS_RING_GEOM")) {
      // threads x lanes per row of the ring kernel: 128x4, 64x4 or 64x2
      if (!std::strcmp(v, "64x4")) { op->ring_nt = 64; op->ring_lpr = 4; }
      else if (!std::strcmp(v, "64x2")) { op->ring_nt = 64; op->ring_lpr = 2; }
      else if (!std::strcmp(v, "128x4")) { op->ring_nt = 128; op->ring_lpr = 4; }
    }
    if (const char* v = std::getenv("PGS_TRSV_GROUPS")) {
      const int g = std::atoi(v);
      if (g >= 0 && g <= 2) op->trsv_groups = g;
    }
    if (const char* v = std::getenv("PGS_TRSV_CG")) {
      const int cg = std::atoi(v);
      if (cg == 1 || cg == 2 || cg == 4 || cg == 8) op->trsv_cg = cg;
    }
    op->is_complex = h->is_complex;
    if (h->is_complex)
      upload_op<cplx<double>, cplx<float>>(op.get(), *h->op_z, prec_mask);
    else
      upload_op<double, float>(op.get(), *h->op_d, prec_mask);
    HIPCHK(hipDeviceSynchronize());
  } catch (const HipError& e) {
    return set_err(e);
  } catch (const std::bad_alloc&) {
    return PGS_ERR_OOM;
  }
  *out = reinterpret_cast<pgs_dev_op_t>(op.release());
  return 0;
}

int pgs_dev_op_destroy(pgs_dev_op_t dop) {
  if (!dop) return 0;
  pgs_dev_op_s* op = reinterpret_cast<pgs_dev_op_s*>(dop);
  (void)hipSetDevice(op->device);
  delete op;
  return 0;
}

int pgs_dev_op_bytes(pgs_dev_op_t dop, int64_t* bytes) {
  if (!dop || !bytes) return PGS_ERROR_INVALID_PARAM;
  *bytes = reinterpret_cast<pgs_dev_op_s*>(dop)->bytes;
  return 0;
}

int pgs_dev_op_dims(pgs_dev_op_t dop, int* n, int* device, int* is_complex) {
  if (!dop || !n || !device || !is_complex) return PGS_ERROR_INVALID_PARAM;
  *n = reinterpret_cast<pgs_dev_op_s*>(dop)->n;
  *device = reinterpret_cast<pgs_dev_op_s*>(dop)->device;
  *is_complex = reinterpret_cast<pgs_dev_op_s*>(dop)->is_complex;
  return 0;
}

int pgs_ws_estimate(pgs_dev_op_t dop, int max_nrhs, int kdim, int64_t* bytes) {
  if (!dop || !bytes || max_nrhs < 1 || max_nrhs > 64 || kdim < 1 || kdim + 1 > kPMAX) return PGS_ERROR_INVALID_PARAM;
  const pgs_dev_op_s* op = reinterpret_cast<pgs_dev_op_s*>(dop);
  int nr = 1;
  while (nr < max_nrhs) nr <<= 1;
  *bytes = (int64_t)ws_sizes(op, nr, kdim).total() + 16;  // + the pinned flag
  return 0;
}

namespace {
int ws_create(pgs_dev_op_s* op, int max_nrhs, int kdim, bool lowrank_build, pgs_ws_t* out) {
  *out = nullptr;
  int nr = 1;
  while (nr < max_nrhs) nr <<= 1;
  std::unique_ptr<pgs_ws_s> ws(new pgs_ws_s());
  ws->op = op;
  ws->max_nrhs = nr;
  ws->kdim = kdim;
  try {
    HIPCHK(hipSetDevice(op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    const WsSizes z = ws_sizes(op, nr, kdim, lowrank_build);
    ws->part_blocks = z.part_blocks;
    ws->v_bytes = (int64_t)z.vb;
    if (const char* v = std::getenv("PGS_FUSED_CGS2")) ws->fused_cgs2 = std::atoi(v) != 0;
    ws->basis_pad = kBasisPad;
    if (const char* v = std::getenv("PGS_BASIS_PAD")) ws->basis_pad = std::atoi(v) == 0 ? 0 : kBasisPad;
    ws->zlev.clear();
    for (size_t b : z.zlev) {
      std::unique_ptr<DevMem> m(new DevMem());
      m->alloc(b);
      ws->zlev.push_back(std::move(m));
    }
    for (const auto& e : z.bufs) (ws.get()->*(e.first)).alloc(e.second);
    HIPCHK(hipHostMalloc((void**)&ws->flag_host, 16, hipHostMallocDefault));
  } catch (const HipError& e) {
    return set_err(e);
  }
  *out = reinterpret_cast<pgs_ws_t>(ws.release());
  return 0;
}
}  // namespace

int pgs_ws_create(pgs_dev_op_t dop, int max_nrhs, int kdim, pgs_ws_t* out) {
  if (!dop || !out || max_nrhs < 1 || max_nrhs > 64 || kdim < 1 || kdim + 1 > kPMAX) return PGS_ERROR_INVALID_PARAM;
  return ws_create(reinterpret_cast<pgs_dev_op_s*>(dop), max_nrhs, kdim, false, out);
}

int pgs_ws_destroy(pgs_ws_t w) {
  if (!w) return 0;
  pgs_ws_s* ws = reinterpret_cast<pgs_ws_s*>(w);
  (void)hipSetDevice(ws->op->device);
  if (ws->flag_host) (void)hipHostFree(ws->flag_host);
  delete ws;
  return 0;
}

int pgs_spmv(pgs_dev_op_t dop, pgs_ws_t w, const double* alpha, const double* x, const double* beta, double* y,
             int nrhs, void* stream) {
  if (!dop || !x || !y || !alpha || !beta || !pow2(nrhs) || nrhs > 64) return PGS_ERROR_INVALID_PARAM;
  pgs_dev_op_s* op = reinterpret_cast<pgs_dev_op_s*>(dop);
  (void)w;
  try {
    HIPCHK(hipSetDevice(op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    if (op->is_complex) {
      using Z = cplx<double>;
      launch_spmv<Z, Z, Z, Z>(op->A, reinterpret_cast<const Z*>(x), reinterpret_cast<Z*>(y), (const Z*)nullptr,
                              Z(alpha[0], alpha[1]), Z(beta[0], beta[1]), 0, nrhs, (hipStream_t)stream);
    } else {
      launch_spmv<double, double, double, double>(op->A, x, y, (const double*)nullptr, alpha[0], beta[0], 0, nrhs,
                                                  (hipStream_t)stream);
    }
  } catch (const HipError& e) {
    return set_err(e);
  }
  return 0;
}

}  // extern "C"

namespace {

// z = M^{-1} v in the precision `prec` selects, with the apply buffers of ws
int apply_dispatch(pgs_dev_op_s* op, pgs_ws_s* ws, const double* v, double* z, int nrhs, int prec, void* stream) {
  try {
    HIPCHK(hipSetDevice(op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    hipStream_t s = (hipStream_t)stream;
    using Z = cplx<double>;
    const Z* vz = reinterpret_cast<const Z*>(v);
    Z* zz = reinterpret_cast<Z*>(z);
    if (prec == PGS_PREC_FP64) {
      if (op->is_complex) {
        if (!op->pz64.present) return PGS_ERROR_INVALID_OPTION;
        apply_impl<Z, Z, Z>(op, ws, vz, zz, nrhs, s);
      } else {
        if (!op->p64.present) return PGS_ERROR_INVALID_OPTION;
        apply_impl<double, double, double>(op, ws, v, z, nrhs, s);
      }
    } else if (prec == PGS_PREC_FP32) {
      if (op->is_complex) {
        if (!op->pz32.present) return PGS_ERROR_INVALID_OPTION;
        apply_impl<cplx<float>, Z, Z>(op, ws, vz, zz, nrhs, s);
      } else {
        if (!op->p32.present) return PGS_ERROR_INVALID_OPTION;
        apply_impl<float, double, double>(op, ws, v, z, nrhs, s);
      }
    } else {
      return PGS_ERROR_INVALID_OPTION;
    }
  } catch (const HipError& e) {
    return set_err(e);
  } catch (const std::exception&) {
    return PGS_ERROR_FUNCTION_CALL;
  }
  return 0;
}

}  // namespace

extern "C" {

int pgs_apply(pgs_dev_op_t dop, pgs_ws_t w, const double* v, double* z, int nrhs, int prec, void* stream) {
  if (!dop || !w || !v || !z) return PGS_ERROR_INVALID_PARAM;
  pgs_dev_op_s* op = reinterpret_cast<pgs_dev_op_s*>(dop);
  pgs_ws_s* ws = reinterpret_cast<pgs_ws_s*>(w);
  if (ws->op != op || !pow2(nrhs) || nrhs > ws->max_nrhs) return PGS_ERROR_INVALID_PARAM;
  return apply_dispatch(op, ws, v, z, nrhs, prec, stream);
}

}  // extern "C"

namespace {

template <int PM, int MODE, int LPE>
void launch_basis_mode(int lds_max, int nblk, size_t lds, int64_t n, int nrhs, int p, const double* V, int64_t ldv,
                       const double* h, double* w, double* part, hipStream_t s) {
  (void)lds_max;  // LDS limit raised in set_kernel_attrs
  k_basis_reg<kBasisNT, PM, MODE, LPE><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, h, w, part);
}

size_t basis_lds(int p, int nrhs, int mode) {
  const int q = mode == 2 ? 1 : p;
  return sizeof(double) * ((mode ? (size_t)p * nrhs : 0) + (size_t)(kBasisNT / kWave) * q * nrhs);
}

// k_basis_reg takes p basis vectors of nrhs columns in every mode
bool basis_fits(const pgs_ws_s* ws, int p, int nrhs) {
  return p <= kBasisPM && basis_lds(p, nrhs, 1) <= (size_t)ws->op->lds_max;
}

// One sweep of k_basis_reg over V_0..V_{p-1} (basis_fits must hold): mode 0
// out[j, c] = V_j^T w; mode 1 w -= V h, then out[j, c] = V_j^T w; mode 2
// w -= V h, then out[c] = ||w(:, c)||^2.  Past 32 vectors the dot sweeps
// spread an element over two lanes (32 basis values each) to keep register
// use, and with it occupancy, at the level of the 32-vector case.
void basis_reg(pgs_ws_s* ws, int64_t n, int nrhs, int p, const double* V, int64_t ldv, const double* h, double* w,
               double* out, int mode, hipStream_t s) {
  const int q = mode == 2 ? 1 : p;
  const size_t lds = basis_lds(p, nrhs, mode);
  const int lds_max = ws->op->lds_max;
  const bool two = p > 32 && mode != 2 && nrhs <= 32;
  const int64_t nthr = (n * nrhs * (two ? 2 : 1) + kBasisNT - 1) / kBasisNT;
  const int nblk = (int)std::min<int64_t>(ws->part_blocks, nthr);
  double* part = ws->part.as<double>();
#define PGS_BASIS(PM, LPE)                                                                                   \
  do {                                                                                                       \
    if (mode == 0) launch_basis_mode<PM, 0, LPE>(lds_max, nblk, lds, n, nrhs, p, V, ldv, h, w, part, s);     \
    else if (mode == 1) launch_basis_mode<PM, 1, LPE>(lds_max, nblk, lds, n, nrhs, p, V, ldv, h, w, part, s); \
    else launch_basis_mode<PM, 2, 1>(lds_max, nblk, lds, n, nrhs, p, V, ldv, h, w, part, s);                \
  } while (0)
  if (p <= 8) PGS_BASIS(8, 1);
  else if (p <= 16) PGS_BASIS(16, 1);
  else if (p <= 32) PGS_BASIS(32, 1);
  else if (two) PGS_BASIS(32, 2);
  else PGS_BASIS(64, 1);
#undef PGS_BASIS
  k_dot_final_wave<<<(q * nrhs + 3) / 4, 256, 0, s>>>(nblk, q, nrhs, part, out, 0);
  HIPCHK(hipGetLastError());
}

// apply arithmetic of the two preconditioner precisions for Krylov scalar S
template <typename S>
struct PrecTypes {
  using P64 = double;
  using P32 = float;
};
template <>
struct PrecTypes<cplx<double>> {
  using P64 = cplx<double>;
  using P32 = cplx<float>;
};

// FGMRES (FlexGmresClass::Solve, fgmres.hpp:300-540) on nrhs columns at once:
// Krylov scalar S (double or complex<double>), preconditioned basis Z in VZ.
// Delayed CGS2, dots of step p: d = [s (p rows) | a (p rows) | q.q | q.w]
// with s = V_{0..p-1}^H q, a = V_{0..p-1}^H w, q = basis slot p; with w ==
// nullptr (the closing pass of a cycle) d = [s | q.q].  The basis is read in
// chunks of 16 vectors per lane for real values, 8 for complex ones (two
// lanes per element up to 32 columns).
template <typename S>
void dcgs2_dots(pgs_ws_s* ws, int64_t n, int nrhs, int p, const S* V, int64_t ldv, const S* w, S* d,
                hipStream_t s) {
  constexpr int PM = scalar_traits<S>::is_complex ? 8 : 16;
  const S* q = V + (int64_t)p * ldv;
  // two lanes per element up to 32 columns; at 64 columns one lane per
  // element runs faster although its chunks of 16 vectors read q and w twice
  // as often (128^3, 64 columns: 478 against 500 ms of dots per two solves)
  const int lpe = nrhs <= 32 ? 2 : 1;
  const int chunk = PM * lpe;
  const int64_t nthr = (n * nrhs * lpe + kBasisNT - 1) / kBasisNT;
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(ws->part_blocks, nthr));
  const int hw = w != nullptr;
  S* part = ws->part.as<S>();
  for (int j0 = 0; j0 == 0 || j0 < p; j0 += chunk) {
    const int cnt = std::min(chunk, p - j0);
    const int first = j0 == 0;
    const int nout = (hw ? 2 * cnt : cnt) + (first ? (hw ? 2 : 1) : 0);
    const size_t lds = sizeof(S) * (size_t)nout * nrhs;
    if (lpe == 2)
      k_dcgs2_dots<S, kBasisNT, PM, 2><<<nblk, kBasisNT, lds, s>>>(n, nrhs, j0, cnt, V, ldv, q, w, first, part);
    else
      k_dcgs2_dots<S, kBasisNT, PM, 1><<<nblk, kBasisNT, lds, s>>>(n, nrhs, j0, cnt, V, ldv, q, w, first, part);
    k_dcgs2_final<S><<<(nout * nrhs + 3) / 4, 256, 0, s>>>(nblk, nout, cnt, nrhs, hw, p, j0, part, d);
  }
  HIPCHK(hipGetLastError());
}

// Delayed CGS2, update of step p (k_dcgs2_update: w minus its projection on
// the final basis v_0..v_p, in terms of the raw slots 0..p), then
// t2[c] = ||w(:, c)||^2.
// Real values: one lane per element up to 32 vectors, then two (the one-lane
// 64-vector kernel needs more than 256 VGPRs); complex values (twice the
// registers per vector): one lane up to 16 vectors, then two, at most 64.
template <typename S>
void dcgs2_update(pgs_ws_s* ws, int64_t n, int nrhs, int p, const S* V, int64_t ldv, const S* q, S* w,
                  const S* coef, double* t2, hipStream_t s) {
  constexpr bool cx = scalar_traits<S>::is_complex;
  const bool two = cx ? p > 16 : p > 32;
  const int64_t nthr = (n * nrhs * (two ? 2 : 1) + kBasisNT - 1) / kBasisNT;
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(ws->part_blocks, nthr));
  const size_t lds = sizeof(S) * (size_t)(p + 1) * nrhs + sizeof(double) * (size_t)(kBasisNT / kWave) * nrhs;
  double* part = ws->part.as<double>();
  if constexpr (cx) {
    if (p <= 8) k_dcgs2_update<S, kBasisNT, 8, 1><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else if (p <= 16) k_dcgs2_update<S, kBasisNT, 16, 1><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else if (p <= 32) k_dcgs2_update<S, kBasisNT, 16, 2><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else if (p <= 64) k_dcgs2_update<S, kBasisNT, 32, 2><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else throw std::runtime_error("delayed CGS2: complex bases hold at most 64 vectors");
  } else {
    if (p <= 8) k_dcgs2_update<S, kBasisNT, 8, 1><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else if (p <= 16) k_dcgs2_update<S, kBasisNT, 16, 1><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else if (p <= 32) k_dcgs2_update<S, kBasisNT, 32, 1><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else if (p <= 64) k_dcgs2_update<S, kBasisNT, 32, 2><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
    else k_dcgs2_update<S, kBasisNT, 64, 2><<<nblk, kBasisNT, lds, s>>>(n, nrhs, p, V, ldv, q, w, coef, part);
  }
  k_dot_final_wave<<<(nrhs + 3) / 4, 256, 0, s>>>(nblk, 1, nrhs, part, t2, 0);
  HIPCHK(hipGetLastError());
}

template <typename S, typename VZ>
int fgmres_impl(pgs_dev_op_s* op, pgs_ws_s* ws, const S* b, S* x, int nrhs, const pgs_fgmres_opts_t& o,
                pgs_fgmres_result_t* res, hipStream_t s) {
  const int64_t n = op->n;
  const int m = o.kdim;
  const int64_t ldv = n * nrhs + ws->basis_pad;
  S* V = ws->V.as<S>();
  VZ* Z = ws->Z.as<VZ>();
  KState* st = ws->st.as<KState>();
  S* H = ws->H.as<S>();
  S* rs = ws->rs.as<S>();
  const int nb = nblocks_for(n * nrhs, kNT);
  const int check = o.check_every > 0 ? o.check_every : 1;
  const int maxits = o.maxits;
  // delayed CGS2 keeps a step's coefficients in LDS and its basis values in
  // registers: real bases up to the Krylov limit (two lanes per element past
  // 32 vectors), complex bases up to 64 vectors (CGS2 beyond)
  const bool dcgs2 = o.orth == PGS_ORTH_DCGS2 && (!scalar_traits<S>::is_complex || m <= 64);
  // ||b||, v0 = b - A x, ||v0||
  colnorm2<S>(ws, n, nrhs, b, ws->nb2.as<double>(), s);
  timed(ws, s, PGS_PROF_SPMV_A, spmv_bytes<S, S, S>(op->A, nrhs, true),
        [&] { launch_spmv<S, S, S, S>(op->A, x, V, b, S(1.0), S(0.0), 1, nrhs, s); });
  colnorm2<S>(ws, n, nrhs, V, ws->nr2.as<double>(), s);
  k_fgmres_init<<<1, 64, 0, s>>>(nrhs, ws->nb2.as<double>(), ws->nr2.as<double>(), o.tol, o.absolute_tol, st,
                                 ws->zflag.as<double>());
  k_zero_cols<S><<<nb, kNT, 0, s>>>(n, nrhs, ws->zflag.as<double>(), x);
  HIPCHK(hipGetLastError());
  int* hflag = ws->flag_host;
  // number of running (restartable) or active columns, read back through pinned memory
  auto read_flag = [&](bool running) -> int {
    if (running)
      k_count_running<<<1, 1, 0, s>>>(nrhs, st, maxits, ws->flag.as<int>());
    else
      k_count_active<<<1, 1, 0, s>>>(nrhs, st, ws->flag.as<int>());
    HIPCHK(hipMemcpyAsync(hflag, ws->flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return hflag[0];
  };
  // inner step i of a cycle: z = M^{-1} v_{i-1}, w = A z, orthogonalize w
  // against v_0..v_{i-1}, Givens update, v_i = w / ||w||; every launch on
  // stream `ss`, nothing else (it may be captured into a graph)
  auto inner_step = [&](int i, hipStream_t ss) {
    S* vprev = V + (int64_t)(i - 1) * ldv;
    S* w = V + (int64_t)i * ldv;
    VZ* zc = Z + (int64_t)(i - 1) * ldv;
    // delayed CGS2 leaves each new basis vector unnormalized: its readers
    // (this apply's gather, the next sweep's dot table) apply the scale
    const double* vsc = dcgs2 ? ws->dc_vscale.as<double>() : nullptr;
    if (o.use_precond) {
      if (o.prec == PGS_PREC_FP64)
        apply_impl<typename PrecTypes<S>::P64, S, VZ>(op, ws, vprev, zc, nrhs, ss, vsc);
      else
        apply_impl<typename PrecTypes<S>::P32, S, VZ>(op, ws, vprev, zc, nrhs, ss, vsc);
    } else if (vsc) {
      k_copy_scale_cols<S, VZ><<<nb, kNT, 0, ss>>>(n, nrhs, vsc, vprev, zc);
    } else {
      k_copy_conv<S, VZ><<<nb, kNT, 0, ss>>>(n * nrhs, vprev, zc);
    }
    timed(ws, ss, PGS_PROF_SPMV_A, spmv_bytes<S, VZ, S>(op->A, nrhs, false),
          [&] { launch_spmv<S, VZ, S, S>(op->A, zc, w, (const S*)nullptr, S(1.0), S(0.0), 0, nrhs, ss); });
    S* h = ws->hbuf.as<S>();
    S* h2 = ws->h2buf.as<S>();
    double* t2 = ws->t2.as<double>();
    if (dcgs2) {
      // delayed CGS2: finalize v_{i-1} (and Hessenberg column i-2, with its
      // Givens update) from the dots of this sweep, then project w once.
      // The slots keep the raw vectors; the final basis is carried by R
      // (k_dcgs2_fix), so the update writes w alone
      const int p = i - 1;
      S* d = ws->dc_d.as<S>();
      S* coef = ws->dc_coef.as<S>();
      const double vb = (double)n * nrhs * sizeof(S);
      timed(ws, ss, PGS_PROF_ORTH_DOTS, (p + 2.0) * vb, [&] { dcgs2_dots<S>(ws, n, nrhs, p, V, ldv, w, d, ss); });
      timed(ws, ss, PGS_PROF_STATE, 0.0, [&] {
        k_dcgs2_fix<S><<<nrhs, kWave, sizeof(S) * 2 * std::max(p, 1), ss>>>(
            nrhs, m, p, maxits, 0, d, H, ws->cs.as<S>(), ws->sn.as<S>(), rs, st, coef, vsc, ws->dc_R.as<S>());
      });
      timed(ws, ss, PGS_PROF_ORTH_UPDATE, (p + 3.0) * vb,
            [&] { dcgs2_update<S>(ws, n, nrhs, p, V, ldv, vprev, w, coef, t2, ss); });
      timed(ws, ss, PGS_PROF_STATE, 0.0,
            [&] { k_dcgs2_norm<S><<<1, 64, 0, ss>>>(nrhs, m, p, t2, H, st, ws->dc_vscale.as<double>()); });
      HIPCHK(hipGetLastError());
      return;
    }
    if (o.orth == PGS_ORTH_CGS2 || o.orth == PGS_ORTH_DCGS2) {
      // PargemslrCgs2 (matrixops.cpp:10588-10709): two projection passes, the
      // norm of the result accumulated by the second update.  Three sweeps
      // over the basis: h = V^T w; w -= V h fused with h2 = V^T w; w -= V h2.
      bool fused = false;
      if constexpr (!scalar_traits<S>::is_complex) {
        if (ws->fused_cgs2 && basis_fits(ws, i, nrhs)) {
          const double vb = (double)n * nrhs * sizeof(S);
          timed(ws, ss, PGS_PROF_ORTH_DOTS, (i + 1.0) * vb,
                [&] { basis_reg(ws, n, nrhs, i, V, ldv, nullptr, w, h, 0, ss); });
          timed(ws, ss, PGS_PROF_ORTH_UPDATE, (i + 2.0) * vb,
                [&] { basis_reg(ws, n, nrhs, i, V, ldv, h, w, h2, 1, ss); });
          timed(ws, ss, PGS_PROF_ORTH_UPDATE, (i + 2.0) * vb,
                [&] { basis_reg(ws, n, nrhs, i, V, ldv, h2, w, t2, 2, ss); });
          fused = true;
        }
      }
      if (!fused) {
        const double vb = (double)n * nrhs * sizeof(S);
        timed(ws, ss, PGS_PROF_ORTH_DOTS, (i + 1.0) * vb, [&] { multidot<S, S, S>(ws, n, nrhs, i, V, ldv, w, h, 0, ss); });
        timed(ws, ss, PGS_PROF_ORTH_UPDATE, (i + 2.0) * vb, [&] { multi_axpy<S>(ws, n, nrhs, i, V, ldv, h, w, nullptr, ss); });
        timed(ws, ss, PGS_PROF_ORTH_DOTS, (i + 1.0) * vb, [&] { multidot<S, S, S>(ws, n, nrhs, i, V, ldv, w, h2, 0, ss); });
        timed(ws, ss, PGS_PROF_ORTH_UPDATE, (i + 2.0) * vb, [&] { multi_axpy<S>(ws, n, nrhs, i, V, ldv, h2, w, t2, ss); });
      }
      dot_final<S>(1, i, nrhs, h2, h, 1, ss);
    } else {
      // PargemslrMgs (matrixops.cpp:10728-10807): one (dot, axpy) pair per basis vector
      const double vb = (double)n * nrhs * sizeof(S);
      for (int j = 0; j < i; ++j) {
        timed(ws, ss, PGS_PROF_ORTH_DOTS, 2.0 * vb, [&] {
          multidot<S, S, S>(ws, n, nrhs, 1, V + (int64_t)j * ldv, ldv, w, h + (int64_t)j * nrhs, 0, ss);
        });
        timed(ws, ss, PGS_PROF_ORTH_UPDATE, 3.0 * vb, [&] {
          multi_axpy<S>(ws, n, nrhs, 1, V + (int64_t)j * ldv, ldv, h + (int64_t)j * nrhs, w,
                        j == i - 1 ? t2 : nullptr, ss);
        });
      }
    }
    timed(ws, ss, PGS_PROF_STATE, 0.0, [&] {
      k_givens<S><<<1, 64, 0, ss>>>(nrhs, m, i, maxits, h, ws->t2.as<double>(), H, ws->cs.as<S>(), ws->sn.as<S>(),
                                    rs, st);
    });
    timed(ws, ss, PGS_PROF_SCALE, 2.0 * (double)n * nrhs * sizeof(S),
          [&] { k_scale_cols<S><<<nb, kNT, 0, ss>>>(n, nrhs, st, w); });
    HIPCHK(hipGetLastError());
  };
  // graphs of the inner steps, kept per workspace while the settings that
  // enter the recorded launches stay the same (not while timing kernels)
  const bool graphs = o.use_graph != 0 && !ws->prof;
  if (graphs) {
    const int key[8] = {nrhs, o.prec, o.orth, o.use_precond, maxits, m, ws->fused_cgs2 ? 1 : 0,
                        (int)sizeof(VZ)};
    if (!ws->cap_stream) HIPCHK(hipStreamCreateWithFlags(&ws->cap_stream, hipStreamNonBlocking));
    if (std::memcmp(key, ws->graph_key, sizeof(key)) != 0 || (int)ws->graph_exec.size() != m + 1) {
      ws->drop_graphs();
      ws->graph_exec.assign(m + 1, nullptr);
      std::memcpy(ws->graph_key, key, sizeof(key));
    }
  }
  int iters_host = 0;
  while (true) {
    if (read_flag(true) == 0) break;
    k_cycle_start<S><<<1, 64, 0, s>>>(nrhs, m, st, rs, dcgs2 ? ws->dc_vscale.as<double>() : nullptr);
    k_scale_cols<S><<<nb, kNT, 0, s>>>(n, nrhs, st, V);
    int i = 0;
    bool none_active = false;
    while (i < m && iters_host < maxits) {
      ++i;
      ++iters_host;
      if (graphs) {
        hipGraphExec_t& ex = ws->graph_exec[i];
        if (!ex) {
          // record inner step i once on the workspace's capture stream
          HIPCHK(hipStreamBeginCapture(ws->cap_stream, hipStreamCaptureModeThreadLocal));
          hipGraph_t g = nullptr;
          try {
            inner_step(i, ws->cap_stream);
          } catch (...) {
            (void)hipStreamEndCapture(ws->cap_stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
          }
          HIPCHK(hipStreamEndCapture(ws->cap_stream, &g));
          const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          if (e != hipSuccess) {
            ex = nullptr;
            throw HipError{e};
          }
        }
        HIPCHK(hipGraphLaunch(ex, s));
      } else {
        inner_step(i, s);
      }
      if (i % check == 0 || i == m) {
        if (read_flag(false) == 0) {
          none_active = true;
          break;
        }
      }
    }
    if (dcgs2 && !none_active) {
      // close the cycle: finalize the last basis vector and Hessenberg
      // column of the columns still running (one more dot sweep)
      S* d = ws->dc_d.as<S>();
      timed(ws, s, PGS_PROF_ORTH_DOTS, (i + 1.0) * (double)n * nrhs * sizeof(S),
            [&] { dcgs2_dots<S>(ws, n, nrhs, i, V, ldv, nullptr, d, s); });
      timed(ws, s, PGS_PROF_STATE, 0.0, [&] {
        k_dcgs2_fix<S><<<nrhs, kWave, sizeof(S) * 2 * std::max(i, 1), s>>>(
            nrhs, m, i, maxits, 1, d, H, ws->cs.as<S>(), ws->sn.as<S>(), rs, st, ws->dc_coef.as<S>(),
            ws->dc_vscale.as<double>(), ws->dc_R.as<S>());
      });
      HIPCHK(hipGetLastError());
    }
    // x += Z y
    timed(ws, s, PGS_PROF_STATE, 0.0,
          [&] { k_backsolve<S><<<1, 64, 0, s>>>(nrhs, m, H, rs, st, ws->coef.as<S>(), ws->upd.as<int>()); });
    timed(ws, s, PGS_PROF_UPDATE_X, (double)n * nrhs * ((double)i * sizeof(VZ) + 2.0 * sizeof(S)), [&] {
      k_update_x<VZ, S><<<nb, kNT, sizeof(S) * std::max(1, i) * nrhs, s>>>(n, nrhs, i, Z, ldv, ws->coef.as<S>(),
                                                                          ws->upd.as<int>(), x);
    });
    HIPCHK(hipGetLastError());
    if (iters_host >= maxits) break;
    if (read_flag(true) == 0) break;
    // restart: v0 = b - A x (fgmres.hpp:500-514); normr keeps the Givens estimate
    timed(ws, s, PGS_PROF_SPMV_A, spmv_bytes<S, S, S>(op->A, nrhs, true),
          [&] { launch_spmv<S, S, S, S>(op->A, x, V, b, S(1.0), S(0.0), 1, nrhs, s); });
  }
  std::vector<KState> hs(nrhs);
  HIPCHK(hipMemcpyAsync(hs.data(), st, sizeof(KState) * nrhs, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int c = 0; c < nrhs; ++c) {
    res[c].iters = hs[c].iters;
    res[c].rel_res = hs[c].normb > 0 ? hs[c].normr / hs[c].normb : 0.0;
    if (hs[c].status == 1 || hs[c].status == 4) res[c].status = 0;
    else if (hs[c].status == 2) res[c].status = PGS_ERR_BREAKDOWN;
    else if (hs[c].status == 5) res[c].status = PGS_ERR_NONFINITE;
    else res[c].status = PGS_ERR_NOT_CONVERGED;
    if (hs[c].status == 4) res[c].rel_res = 0.0;
  }
  return 0;
}

}  // namespace

extern "C" int pgs_fgmres(pgs_dev_op_t dop, pgs_ws_t w, const double* b, double* x, int nrhs,
                          const pgs_fgmres_opts_t* opts, pgs_fgmres_result_t* res, void* stream) {
  if (!dop || !w || !b || !x || !opts || !res) return PGS_ERROR_INVALID_PARAM;
  pgs_dev_op_s* op = reinterpret_cast<pgs_dev_op_s*>(dop);
  pgs_ws_s* ws = reinterpret_cast<pgs_ws_s*>(w);
  if (ws->op != op || !pow2(nrhs) || nrhs > ws->max_nrhs) return PGS_ERROR_INVALID_PARAM;
  if (opts->kdim < 1 || opts->kdim > ws->kdim || opts->maxits < 0) return PGS_ERROR_INVALID_PARAM;
  if (opts->orth != PGS_ORTH_MGS && opts->orth != PGS_ORTH_CGS2 && opts->orth != PGS_ORTH_DCGS2)
    return PGS_ERROR_INVALID_OPTION;
  if (opts->use_precond && opts->prec != PGS_PREC_FP64 && opts->prec != PGS_PREC_FP32) return PGS_ERROR_INVALID_OPTION;
  const bool has64 = op->is_complex ? op->pz64.present : op->p64.present;
  const bool has32 = op->is_complex ? op->pz32.present : op->p32.present;
  if (opts->use_precond && opts->prec == PGS_PREC_FP64 && !has64) return PGS_ERROR_INVALID_OPTION;
  if (opts->use_precond && opts->prec == PGS_PREC_FP32 && !has32) return PGS_ERROR_INVALID_OPTION;
  try {
    HIPCHK(hipSetDevice(op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    hipStream_t s = (hipStream_t)stream;
    const bool low = opts->use_precond && opts->prec == PGS_PREC_FP32;
    if (op->is_complex) {
      using Z = cplx<double>;
      const Z* bz = reinterpret_cast<const Z*>(b);
      Z* xz = reinterpret_cast<Z*>(x);
      if (low) return fgmres_impl<Z, cplx<float>>(op, ws, bz, xz, nrhs, *opts, res, s);
      return fgmres_impl<Z, Z>(op, ws, bz, xz, nrhs, *opts, res, s);
    }
    if (low) return fgmres_impl<double, float>(op, ws, b, x, nrhs, *opts, res, s);
    return fgmres_impl<double, double>(op, ws, b, x, nrhs, *opts, res, s);
  } catch (const HipError& e) {
    return set_err(e);
  } catch (const std::exception&) {
    return PGS_ERROR_FUNCTION_CALL;
  }
}

extern "C" const char* pgs_last_hip_error(void) { return g_last_hip_error.c_str(); }

namespace {
int hip_rc(hipError_t e) {
  if (e == hipSuccess) return 0;
  return set_err(HipError{e});
}
}  // namespace

extern "C" int pgs_dev_malloc(int device, int64_t bytes, void** ptr) {
  if (!ptr || bytes < 0) return PGS_ERROR_INVALID_PARAM;
  *ptr = nullptr;
  int rc = hip_rc(hipSetDevice(device));
  if (rc) return rc;
  return hip_rc(hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1)));
}

extern "C" int pgs_dev_free(int device, void* ptr) {
  if (!ptr) return 0;
  int rc = hip_rc(hipSetDevice(device));
  if (rc) return rc;
  return hip_rc(hipFree(ptr));
}

extern "C" int pgs_memcpy_htod(int device, void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return PGS_ERROR_INVALID_PARAM;
  int rc = hip_rc(hipSetDevice(device));
  if (rc || bytes == 0) return rc;
  rc = hip_rc(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  if (rc) return rc;
  return hip_rc(hipStreamSynchronize((hipStream_t)stream));
}

extern "C" int pgs_memcpy_dtoh(int device, void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return PGS_ERROR_INVALID_PARAM;
  int rc = hip_rc(hipSetDevice(device));
  if (rc || bytes == 0) return rc;
  rc = hip_rc(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  if (rc) return rc;
  return hip_rc(hipStreamSynchronize((hipStream_t)stream));
}

extern "C" int pgs_memcpy_dtod(int device, void* dst, const void* src, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return PGS_ERROR_INVALID_PARAM;
  int rc = hip_rc(hipSetDevice(device));
  if (rc || bytes == 0) return rc;
  rc = hip_rc(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  if (rc) return rc;
  return hip_rc(hipStreamSynchronize((hipStream_t)stream));
}

extern "C" int pgs_memset_zero(int device, void* ptr, int64_t bytes, void* stream) {
  if (bytes < 0 || (bytes > 0 && !ptr)) return PGS_ERROR_INVALID_PARAM;
  int rc = hip_rc(hipSetDevice(device));
  if (rc || bytes == 0) return rc;
  return hip_rc(hipMemsetAsync(ptr, 0, (size_t)bytes, (hipStream_t)stream));
}

extern "C" int pgs_stream_create(int device, void** stream) {
  if (!stream) return PGS_ERROR_INVALID_PARAM;
  int rc = hip_rc(hipSetDevice(device));
  if (rc) return rc;
  hipStream_t s = nullptr;
  rc = hip_rc(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return rc;
}

extern "C" int pgs_stream_destroy(int device, void* stream) {
  if (!stream) return 0;
  int rc = hip_rc(hipSetDevice(device));
  if (rc) return rc;
  return hip_rc(hipStreamDestroy((hipStream_t)stream));
}

extern "C" int pgs_stream_synchronize(int device, void* stream) {
  int rc = hip_rc(hipSetDevice(device));
  if (rc) return rc;
  return hip_rc(hipStreamSynchronize((hipStream_t)stream));
}

extern "C" int pgs_device_synchronize(int device) {
  int rc = hip_rc(hipSetDevice(device));
  if (rc) return rc;
  return hip_rc(hipDeviceSynchronize());
}

// ---------------------------------------------------------------------------
// device vectors and stand-alone device matrices of the class surface
// ---------------------------------------------------------------------------
namespace {

constexpr int kVecBlocks = 1024;  // grid cap of the single-vector kernels (grid-stride loops)

unsigned vec_grid(int64_t n) { return (unsigned)std::min<int64_t>(kVecBlocks, std::max<int64_t>(1, (n + 255) / 256)); }

bool vec_args_ok(int64_t n, const void* a, const void* b) { return n >= 0 && (n == 0 || (a && b)); }

}  // namespace

struct pgs_dev_csr_s {
  int device = 0;
  int is_complex = 0;
  DevCsr A;
};

extern "C" int pgs_vec_axpy(int device, int64_t n, int is_complex, const double* alpha, const double* x, double* y,
                            void* stream) {
  if (!alpha || !vec_args_ok(n, x, y)) return PGS_ERROR_INVALID_PARAM;
  try {
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this launch's
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (is_complex) {
      using Z = cplx<double>;
      k_vec_axpy<Z><<<vec_grid(n), 256, 0, s>>>(n, Z(alpha[0], alpha[1]), reinterpret_cast<const Z*>(x),
                                                reinterpret_cast<Z*>(y));
    } else {
      k_vec_axpy<double><<<vec_grid(n), 256, 0, s>>>(n, alpha[0], x, y);
    }
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    return set_err(e);
  }
  return 0;
}

extern "C" int pgs_vec_scale(int device, int64_t n, int is_complex, const double* alpha, double* x, void* stream) {
  if (!alpha || !vec_args_ok(n, x, x)) return PGS_ERROR_INVALID_PARAM;
  try {
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this launch's
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (is_complex) {
      using Z = cplx<double>;
      k_vec_scale<Z><<<vec_grid(n), 256, 0, s>>>(n, Z(alpha[0], alpha[1]), reinterpret_cast<Z*>(x));
    } else {
      k_vec_scale<double><<<vec_grid(n), 256, 0, s>>>(n, alpha[0], x);
    }
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    return set_err(e);
  }
  return 0;
}

extern "C" int pgs_vec_fill(int device, int64_t n, int is_complex, const double* value, double* x, void* stream) {
  if (!value || !vec_args_ok(n, x, x)) return PGS_ERROR_INVALID_PARAM;
  try {
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this launch's
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (is_complex) {
      using Z = cplx<double>;
      k_vec_fill<Z><<<vec_grid(n), 256, 0, s>>>(n, Z(value[0], value[1]), reinterpret_cast<Z*>(x));
    } else {
      k_vec_fill<double><<<vec_grid(n), 256, 0, s>>>(n, value[0], x);
    }
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    return set_err(e);
  }
  return 0;
}

extern "C" int pgs_vec_dot(int device, int64_t n, int is_complex, const double* x, const double* y, double* result,
                           void* stream) {
  if (!result || !vec_args_ok(n, x, y)) return PGS_ERROR_INVALID_PARAM;
  result[0] = 0.0;
  if (is_complex) result[1] = 0.0;
  void* buf = nullptr;
  hipStream_t s = (hipStream_t)stream;
  try {
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this launch's
    if (n == 0) return 0;
    const unsigned g = vec_grid(n);
    // partials (2 per workgroup) and the two sums, stream-ordered so that
    // concurrent calls on other streams never share them
    HIPCHK(hipMallocAsync(&buf, sizeof(double) * (2 * (size_t)g + 2), s));
    double* part = static_cast<double*>(buf);
    double* out = part + 2 * (size_t)g;
    if (is_complex) {
      using Z = cplx<double>;
      k_vec_dot<Z><<<g, 256, 0, s>>>(n, reinterpret_cast<const Z*>(x), reinterpret_cast<const Z*>(y), part);
    } else {
      k_vec_dot<double><<<g, 256, 0, s>>>(n, x, y, part);
    }
    HIPCHK(hipGetLastError());
    k_dot_final_wave<<<1, 256, 0, s>>>((int)g, 2, 1, part, out, 0);
    HIPCHK(hipGetLastError());
    double h[2];
    HIPCHK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipFreeAsync(buf, s));
    buf = nullptr;
    HIPCHK(hipStreamSynchronize(s));
    result[0] = h[0];
    if (is_complex) result[1] = h[1];
  } catch (const HipError& e) {
    if (buf) (void)hipFreeAsync(buf, s);
    return set_err(e);
  }
  return 0;
}

extern "C" int pgs_dev_csr_create(int device, int n, const int* ia, const int* ja, const double* a, int is_complex,
                                  pgs_dev_csr_t* out) {
  if (!out || n < 0 || !ia) return PGS_ERROR_INVALID_PARAM;
  *out = nullptr;
  if (ia[0] != 0) return PGS_ERROR_INVALID_PARAM;
  for (int i = 0; i < n; ++i)
    if (ia[i + 1] < ia[i]) return PGS_ERROR_INVALID_PARAM;
  const int64_t nnz = ia[n];
  if (nnz > 0 && (!ja || !a)) return PGS_ERROR_INVALID_PARAM;
  for (int64_t k = 0; k < nnz; ++k)
    if (ja[k] < 0 || ja[k] >= n) return PGS_ERROR_INVALID_PARAM;
  std::unique_ptr<pgs_dev_csr_s> m;
  try {
    m.reset(new pgs_dev_csr_s);
    m->device = device;
    m->is_complex = is_complex ? 1 : 0;
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    if (is_complex) {
      Csr<zdouble> H;
      H.nrows = H.ncols = n;
      H.ia.assign(ia, ia + n + 1);
      H.ja.assign(ja, ja + nnz);
      H.a.resize((size_t)nnz);
      for (int64_t k = 0; k < nnz; ++k) H.a[k] = zdouble(a[2 * k], a[2 * k + 1]);
      upload_csr<cplx<double>>(m->A, H, nullptr);
    } else {
      Csr<double> H;
      H.nrows = H.ncols = n;
      H.ia.assign(ia, ia + n + 1);
      H.ja.assign(ja, ja + nnz);
      H.a.assign(a, a + nnz);
      upload_csr<double>(m->A, H, nullptr);
    }
  } catch (const HipError& e) {
    return set_err(e);
  } catch (const std::bad_alloc&) {
    return PGS_ERR_OOM;
  }
  *out = m.release();
  return 0;
}

extern "C" int pgs_dev_csr_destroy(pgs_dev_csr_t m) {
  if (!m) return 0;
  (void)hipSetDevice(m->device);
  delete m;
  return 0;
}

extern "C" int pgs_dev_csr_spmv(pgs_dev_csr_t m, const double* alpha, const double* x, const double* beta, double* y,
                                void* stream) {
  if (!m || !alpha || !beta || (m->A.nrows > 0 && (!x || !y))) return PGS_ERROR_INVALID_PARAM;
  try {
    HIPCHK(hipSetDevice(m->device));
    (void)hipGetLastError();  // an earlier call's error is not this launch's
    if (m->is_complex) {
      using Z = cplx<double>;
      launch_spmv<Z, Z, Z, Z>(m->A, reinterpret_cast<const Z*>(x), reinterpret_cast<Z*>(y), (const Z*)nullptr,
                              Z(alpha[0], alpha[1]), Z(beta[0], beta[1]), 0, 1, (hipStream_t)stream);
    } else {
      launch_spmv<double, double, double, double>(m->A, x, y, (const double*)nullptr, alpha[0], beta[0], 0, 1,
                                                  (hipStream_t)stream);
    }
  } catch (const HipError& e) {
    return set_err(e);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// z = M^{-1} v without allocating: the apply buffers live in `scratch`, a
// file-scope device array of PGS_APPLY_NOALLOC_SCRATCH bytes that every device
// the library runs on holds until the process ends
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PGS_APPLY_NOALLOC_SCRATCH];

namespace {

std::mutex g_scratch_m;
std::vector<void*> g_scratch_ptr;  // device address of `scratch`, per device

// The device address of `scratch` on `device` (current), looked up once per
// device.  In host code the array's name is not a device address; only this
// pointer is passed to kernels.
int scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> g(g_scratch_m);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_err(HipError{e != hipSuccess ? e : hipErrorInvalidSymbol});
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

// The apply's buffers (as pgs_ws_create sizes them for one workspace of
// nr columns), placed one after another at 256-byte boundaries; returns the
// bytes they take.
size_t noalloc_layout(const pgs_dev_op_s* op, int nr, pgs_ws_s* ws, unsigned char* base) {
  const WsSizes z = ws_sizes(op, nr, 1);
  size_t off = 0;
  auto place = [&](DevMem& m, size_t b) {
    if (ws) m.view(base + off, b);
    off += (b + 255) / 256 * 256;
  };
  for (size_t b : z.zlev) {
    std::unique_ptr<DevMem> m(new DevMem());
    place(*m, b);
    if (ws) ws->zlev.push_back(std::move(m));
  }
  pgs_ws_s dummy;
  pgs_ws_s* w = ws ? ws : &dummy;
  for (const auto& e : z.bufs) {
    DevMem pgs_ws_s::*f = e.first;
    if (f == &pgs_ws_s::r0 || f == &pgs_ws_s::x0 || f == &pgs_ws_s::lrpart || f == &pgs_ws_s::lrt)
      place(w->*f, e.second);
  }
  return off;
}

}  // namespace

extern "C" int pgs_apply_noalloc(pgs_dev_op_t dop, const double* v, double* z, int nrhs, int prec, void* stream) {
  if (!dop || !v || !z || !pow2(nrhs) || nrhs > 64) return PGS_ERROR_INVALID_PARAM;
  pgs_dev_op_s* op = reinterpret_cast<pgs_dev_op_s*>(dop);
  // the inner Schur iteration needs a Krylov workspace: use pgs_apply
  if (op->inner_maxits > 0) return PGS_ERR_UNSUPPORTED;
  const size_t need = noalloc_layout(op, nrhs, nullptr, nullptr);
  if (need > (size_t)PGS_APPLY_NOALLOC_SCRATCH) {
    g_last_hip_error = "pgs_apply_noalloc: the apply needs " + std::to_string(need) + " bytes of scratch, more than " +
                       std::to_string((size_t)PGS_APPLY_NOALLOC_SCRATCH);
    return PGS_ERR_SCRATCH_LIMIT;
  }
  int rc = hip_rc(hipSetDevice(op->device));
  if (rc) return rc;
  void* base = nullptr;
  rc = scratch_address(op->device, &base);
  if (rc) return rc;
  // a workspace whose apply buffers are views of the scratch array (no
  // allocation: its destructor frees nothing it does not own)
  pgs_ws_s ws;
  ws.op = op;
  ws.max_nrhs = nrhs;
  ws.kdim = 1;
  noalloc_layout(op, nrhs, &ws, static_cast<unsigned char*>(base));
  return apply_dispatch(op, &ws, v, z, nrhs, prec, stream);
}

extern "C" int pgs_ws_profile(pgs_ws_t w, int enable) {
  if (!w) return PGS_ERROR_INVALID_PARAM;
  pgs_ws_s* ws = reinterpret_cast<pgs_ws_s*>(w);
  try {
    HIPCHK(hipSetDevice(ws->op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    prof_flush(ws);
  } catch (const HipError& e) {
    return set_err(e);
  }
  ws->prof = enable != 0;
  for (int f = 0; f < PGS_PROFILE_FAMILIES; ++f) {
    ws->fam_ms[f] = 0.0;
    ws->fam_bytes[f] = 0.0;
    ws->fam_launches[f] = 0;
  }
  return 0;
}

extern "C" int pgs_ws_profile_read(pgs_ws_t w, pgs_kernel_profile_t* out) {
  if (!w || !out) return PGS_ERROR_INVALID_PARAM;
  pgs_ws_s* ws = reinterpret_cast<pgs_ws_s*>(w);
  try {
    HIPCHK(hipSetDevice(ws->op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    prof_flush(ws);
  } catch (const HipError& e) {
    return set_err(e);
  }
  out->launches = ws->fam_launches[PGS_PROF_TRSV0];
  out->total_ms = ws->fam_ms[PGS_PROF_TRSV0];
  out->total_bytes = ws->fam_bytes[PGS_PROF_TRSV0];
  return 0;
}

extern "C" int pgs_ws_profile_families(pgs_ws_t w, pgs_kernel_profile_t* out) {
  if (!w || !out) return PGS_ERROR_INVALID_PARAM;
  pgs_ws_s* ws = reinterpret_cast<pgs_ws_s*>(w);
  try {
    HIPCHK(hipSetDevice(ws->op->device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    prof_flush(ws);
  } catch (const HipError& e) {
    return set_err(e);
  }
  for (int f = 0; f < PGS_PROFILE_FAMILIES; ++f) {
    out[f].launches = ws->fam_launches[f];
    out[f].total_ms = ws->fam_ms[f];
    out[f].total_bytes = ws->fam_bytes[f];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Low-rank build of the setup on the device (GemslrClass::SetupLowRank*,
// gemslr.cpp:2782-3257).  The thick-restart
// Arnoldi driver (Hessenberg matrix, Schur forms, restart choices) runs on
// the host as in the host setup; the basis and every m-length operation live
// on the device: G x = (I - C Ct^{-1}) x + E B^{-1} F Ct^{-1} x with Ct^{-1}
// the device SolveLevel of the levels below (whose corrections were built
// and uploaded first), CGS2 by multidot / multi_axpy, and the restart and
// final bases W = V Y, WH = W R.  The arithmetic is FP64 (complex<double>).

namespace {

// out(:, j) = sum_q X(:, q) Y(q, j), j < p, q < k; X column-major with
// leading dimension m, Y with leading dimension ldy
template <typename V>
__global__ __launch_bounds__(256) void k_lr_combine(int64_t m, int k, int p, const V* __restrict__ X,
                                                    const V* __restrict__ Y, int ldy, V* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int j = 0; j < p; ++j) {
    V s = V(0);
    for (int q = 0; q < k; ++q) s += X[i + q * m] * Y[q + (int64_t)j * ldy];
    out[i + j * m] = s;
  }
}

template <typename V>
__global__ __launch_bounds__(256) void k_lr_divide(int64_t m, V* __restrict__ x, double d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) x[i] = x[i] / V(d);
}

template <typename T, typename V>
struct DevLrBackend final : LrBackend<T> {
  static_assert(sizeof(T) == sizeof(V), "host and device values share their layout");
  static constexpr int kChunk = 64;  // basis vectors per multidot / multi_axpy launch
  pgs_dev_op_s* op;
  pgs_ws_t wsh = nullptr;
  pgs_ws_s* ws = nullptr;
  hipStream_t s = nullptr;
  int l = 0, ncols = 0;
  int64_t m = 0;
  DevMem Vb, Tb, zl, zu, wu, hd, yd, nd;
  explicit DevLrBackend(pgs_dev_op_s* o) : op(o) {}
  ~DevLrBackend() override {
    if (wsh) pgs_ws_destroy(wsh);
  }
  V* col(int j) const { return Vb.as<V>() + (int64_t)j * m; }
  const DevLevel& level() const { return *prec_of<V>(op).levels[l]; }
  void begin(int lev, int mm, int nc) override {
    l = lev;
    m = mm;
    ncols = nc;
    // a workspace of the level temporaries and partial sums only
    const int e = ws_create(op, 1, 1, true, &wsh);
    if (e) throw HipError{e == PGS_ERR_OOM ? hipErrorOutOfMemory : hipErrorUnknown};
    ws = reinterpret_cast<pgs_ws_s*>(wsh);
    Vb.alloc(sizeof(V) * (size_t)m * nc);
    Tb.alloc(sizeof(V) * (size_t)m * nc);
    zl.alloc(sizeof(V) * (size_t)m);
    zu.alloc(sizeof(V) * (size_t)std::max(1, level().nloc));
    wu.alloc(sizeof(V) * (size_t)std::max(1, level().nloc));
    hd.alloc(sizeof(V) * (size_t)nc);
    yd.alloc(sizeof(V) * (size_t)nc * nc);
    nd.alloc(sizeof(double));
  }
  void set_col(int j, const T* x) override {
    HIPCHK(hipMemcpy(col(j), x, sizeof(V) * m, hipMemcpyHostToDevice));
  }
  void apply(int j) override {
    const DevLevel& L = level();
    const V* x = col(j);
    V* y = col(j + 1);
    solve_level<V>(op, ws, prec_of<V>(op), l + 1, x, zl.as<V>(), 1, s);
    launch_spmv<V, V, V, V>(L.C, zl.as<V>(), y, x, V(1.0), V(0.0), 1, 1, s);  // y = x - C zl
    launch_spmv<V, V, V, V>(L.F, zl.as<V>(), zu.as<V>(), (const V*)nullptr, V(1.0), V(0.0), 0, 1, s);
    launch_trsv<V>(op, ws, l, L.B, zu.as<V>(), wu.as<V>(), 1, s);
    launch_spmv<V, V, V, V>(L.E, wu.as<V>(), y, (const V*)nullptr, V(1.0), V(1.0), 0, 1, s);  // y += E wu
  }
  void dots(int j, T* h) override {
    for (int q0 = 0; q0 <= j; q0 += kChunk) {
      const int pc = std::min(kChunk, j + 1 - q0);
      multidot<V, V, V>(ws, m, 1, pc, col(q0), m, col(j + 1), hd.as<V>() + q0, 0, s);
    }
    HIPCHK(hipMemcpy(h, hd.p, sizeof(V) * (size_t)(j + 1), hipMemcpyDeviceToHost));
  }
  void update(int j, const T* h) override {
    HIPCHK(hipMemcpy(hd.p, h, sizeof(V) * (size_t)(j + 1), hipMemcpyHostToDevice));
    for (int q0 = 0; q0 <= j; q0 += kChunk) {
      const int pc = std::min(kChunk, j + 1 - q0);
      multi_axpy<V, V>(ws, m, 1, pc, col(q0), m, hd.as<V>() + q0, col(j + 1), nullptr, s);
    }
  }
  double norm(int j) override {
    colnorm2<V>(ws, m, 1, col(j), nd.as<double>(), s);
    double v = 0.0;
    HIPCHK(hipMemcpy(&v, nd.p, sizeof(double), hipMemcpyDeviceToHost));
    return std::sqrt(v);
  }
  void divide(int j, double d) override {
    k_lr_divide<V><<<nblocks_for(m, 256), 256, 0, s>>>(m, col(j), d);
    HIPCHK(hipGetLastError());
  }
  void combine(const V* X, int k, int p, const T* Y, int ldy, V* out) {
    HIPCHK(hipMemcpy(yd.p, Y, sizeof(V) * (size_t)ldy * p, hipMemcpyHostToDevice));
    k_lr_combine<V><<<nblocks_for(m, 256), 256, 0, s>>>(m, k, p, X, yd.as<V>(), ldy, out);
    HIPCHK(hipGetLastError());
  }
  void restart(int mcur, int p, const T* Y) override {
    combine(Vb.as<V>(), mcur, p, Y, mcur, Tb.as<V>());
    HIPCHK(hipMemcpy(col(p), col(mcur), sizeof(V) * m, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(Vb.p, Tb.p, sizeof(V) * m * p, hipMemcpyDeviceToDevice));
  }
  void finish(int mcur, int r, const T* Y, const T* R, std::vector<T>& W, std::vector<T>& WH) override {
    combine(Vb.as<V>(), mcur, r, Y, mcur, Tb.as<V>());  // W
    combine(Tb.as<V>(), r, r, R, r, Vb.as<V>());       // WH = W R
    W.resize((size_t)m * r);
    WH.resize((size_t)m * r);
    HIPCHK(hipMemcpy(W.data(), Tb.p, sizeof(V) * m * r, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(WH.data(), Vb.p, sizeof(V) * m * r, hipMemcpyDeviceToHost));
  }
};

template <typename T, typename V64, typename V32>
int lowrank_build_device(HostOp<T>& H, int device, std::string* msg) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
    if (msg) *msg = "no such device for the low-rank build";
    return PGS_ERR_NO_DEVICE;
  }
  try {
    std::unique_ptr<pgs_dev_op_s> op(new pgs_dev_op_s());
    HIPCHK(hipSetDevice(device));
    (void)hipGetLastError();  // an earlier call's error is not this call's
    op->device = device;
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    op->lds_max = lds;
    HIPCHK(hipDeviceGetAttribute(&op->num_cus, hipDeviceAttributeMultiprocessorCount, device));
    set_kernel_attrs(lds);
    op->is_complex = scalar_traits<V64>::is_complex ? 1 : 0;
    upload_op<V64, V32>(op.get(), H, PGS_PREC_FP64, true);
    // the setup phase solves with the LU variant of SolveLevel whatever the
    // solve phase will use (gemslr.cpp:3527: the U variant is for kGemslrPhaseSolve)
    op->usolve = 0;
    for (int l = H.nlev - 2; l >= 0; --l) {
      {
        DevLrBackend<T, V64> be(op.get());
        lowrank_level_with(H, l, be);
      }
      // the correction of level l takes part in the operator of level l - 1
      const HostLevel<T>& HL = H.levels[l];
      DevLevel& L = *prec_of<V64>(op.get()).levels[l];
      L.rank = HL.rank;
      if (HL.rank > 0) {
        upload(L.W, to_row_major<V64>(HL.W, L.nrem, HL.rank), &op->bytes);
        upload(L.WH, to_row_major<V64>(HL.WH, L.nrem, HL.rank), &op->bytes);
      }
    }
    HIPCHK(hipDeviceSynchronize());
  } catch (const HipError& e) {
    if (msg) *msg = std::string("device low-rank build: ") + hipGetErrorString(e.e);
    return set_err(e);
  } catch (const std::bad_alloc&) {
    return PGS_ERR_OOM;
  }
  return 0;
}

struct LowrankDeviceInstall {
  LowrankDeviceInstall() {
    g_lowrank_device = &lowrank_build_device<double, double, float>;
    g_lowrank_device_z = &lowrank_build_device<zdouble, cplx<double>, cplx<float>>;
  }
} g_lowrank_device_install;

}  // namespace
