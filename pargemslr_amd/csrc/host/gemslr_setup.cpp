// GeMSLR setup on the host: partition -> level structure -> local RCM ->
// per-block ILUT -> thick-restart Arnoldi low-rank corrections, bottom level
// first.  Mirrors GemslrClass::Setup (ParGeMSLR/SRC/preconditioners/
// gemslr.cpp:1602-1711), SetupPermutationBuildLevelStructure (1907-2211),
// SetupBSolveILUT (2313-2378), SetupLowRank (2507-2585),
// SetupLowRankThickRestartNoLock (2782-2882) and SetupLowRankBuildLowRank
// (3059-3257).
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "dense.hpp"
#include "gemslr_host.hpp"
#include "partition.hpp"

namespace pgs {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline uint64_t mix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline double unif01(uint64_t& s) { return (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0); }

template <typename T> T from_cplx(const dense::cplx& z);
template <> double from_cplx<double>(const dense::cplx& z) { return z.real(); }
template <> zdouble from_cplx<zdouble>(const dense::cplx& z) { return z; }

template <typename T>
void ilu_solve(const IluFactor<T>& F, const T* r, T* x) {
  const int n = F.n;
  for (int i = 0; i < n; ++i) {
    T s = r[i];
    for (int k = F.L.ia[i]; k < F.L.ia[i + 1]; ++k) s -= F.L.a[k] * x[F.L.ja[k]];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    T s = x[i];
    for (int k = F.U.ia[i]; k < F.U.ia[i + 1]; ++k) s -= F.U.a[k] * x[F.U.ja[k]];
    x[i] = s * F.Dinv[i];
  }
}

template <typename T>
void solve_b(const HostOp<T>& op, int l, const T* r, T* x) {
  const HostLevel<T>& L = op.levels[l];
  const int s = op.lev_ptr[l];
  parallel_for(L.ncomp, [&](int b) {
    int off = L.dom_ptr[b] - s;
    ilu_solve(L.ilu[b], r + off, x + off);
  });
}

}  // namespace

void params_default_array(double* p) {
  for (int i = 0; i < 180; ++i) p[i] = 0.0;
  p[1] = 0; p[2] = 1; p[3] = 0; p[4] = 0;
  p[11] = 50; p[12] = 100; p[13] = 1e-6;
  p[20] = 5; p[21] = 5; p[22] = 4; p[23] = 4; p[24] = 1; p[26] = 10;
  p[27] = 1; p[28] = 1; p[29] = 1; p[30] = 1; p[31] = 1; p[32] = 1;
  p[40] = 1e-2; p[41] = 1e-2; p[42] = 1e-2; p[43] = 1e-2; p[44] = 1e-3; p[45] = 1e-3;
  p[46] = 1e-3; p[47] = 1e-3; p[48] = 100; p[49] = 100; p[50] = 200; p[51] = 200;
  p[56] = 200; p[57] = 200; p[58] = 1; p[59] = 1;
  p[70] = 1; p[71] = 1; p[72] = 1; p[73] = 1;
  p[74] = 100; p[75] = 50; p[76] = 100; p[77] = 50;
  p[78] = 1.0; p[79] = 1.0; p[80] = 1.0; p[81] = 1.0;
  p[82] = 2.0; p[83] = 2.0; p[84] = 2.0; p[85] = 2.0;
  p[86] = 1e-9; p[87] = 1e-9; p[88] = 1e-9; p[89] = 1e-9;
  p[90] = 3; p[91] = 3; p[92] = 3; p[93] = 3;
  p[94] = 0.25; p[95] = 1e-14; p[96] = 1.0 / std::sqrt(2.0); p[97] = 1;
  p[130] = 1; p[131] = 1e-2; p[132] = 5;
  p[143] = 1; p[150] = 1; p[160] = 3;
}

void params_from_array(const double* p, GemslrParams& g) {
  g.bsolve_top = (int)p[1];
  g.bsolve_lower = (int)p[3];
  g.global_solve = (int)p[141];
  g.nlev = (int)p[20];
  g.ncomp = (int)p[22];
  g.partition = (int)p[24];
  g.minsep = p[26] > 0 ? (int)p[26] : 10;
  g.kmin = (int)p[27];
  g.kfactor = (int)p[29];
  g.vertexsep = p[31] != 0.0;
  g.droptol_B = p[40];
  g.droptol_C = p[44];
  g.rownnz_B = (int)p[48];
  g.rownnz_C = (int)p[50];
  g.perm_option = (int)p[58];
  g.lr_option1 = (int)p[70];
  g.lr_option2 = (int)p[71];
  g.lr_rank1 = (int)p[74];
  g.lr_rank2 = (int)p[75];
  g.lr_rank_factor1 = p[78];
  g.lr_rank_factor2 = p[79];
  g.lr_arnoldi_factor1 = p[82];
  g.lr_arnoldi_factor2 = p[83];
  g.lr_tol_eig1 = p[86];
  g.lr_tol_eig2 = p[87];
  g.lr_maxits1 = (int)p[90];
  g.lr_maxits2 = (int)p[91];
  g.lr_tr_factor = p[94];
  g.lr_tol_orth = p[95];
  g.lr_rand_init = p[97] != 0.0;
  g.schur_enable = (int)p[130];
  g.schur_tol = p[131];
  g.schur_maxits = (int)p[132];
}

template <typename T>
void host_solve_level(const HostOp<T>& op, int l, const T* rhs, T* x, HostApplyWork<T>& w);

namespace {

template <typename T>
double abs_t(const T& v) { return std::abs(v); }

// y = M (v + W H W^H v) with M the GeMSLR solve of the levels below 0
// (GemslrSchurMatrixClass::Solve, gemslr.cpp:292-320); t: scratch of m values.
template <typename T>
void schur_precond(const HostOp<T>& op, const T* v, T* y, std::vector<T>& t, HostApplyWork<T>& w) {
  const HostLevel<T>& L = op.levels[0];
  const int m = op.nrem(0);
  t.assign(v, v + m);
  if (L.rank > 0) {
    std::vector<T> c(L.rank, T(0));
    for (int j = 0; j < L.rank; ++j) {
      T s = T(0);
      const T* wc = &L.W[(size_t)j * m];
      for (int i = 0; i < m; ++i) s += sconj(wc[i]) * v[i];
      c[j] = s;
    }
    // xlr = WH c, then xlr += v (SolveApplyLowRankLevel + Axpy)
    std::vector<T> xlr(m, T(0));
    for (int j = 0; j < L.rank; ++j) {
      const T* whc = &L.WH[(size_t)j * m];
      for (int i = 0; i < m; ++i) xlr[i] += whc[i] * c[j];
    }
    for (int i = 0; i < m; ++i) t[i] = xlr[i] + v[i];
  }
  host_solve_level(op, 1, t.data(), y, w);
}

// y = S x, S = C - E B^{-1} F on level 0 with B^{-1} the block ILU solve
// (GemslrClass::SchurMatVec, gemslr.cpp:3799-3955, alpha = 1, beta = 0):
// y = -E B^{-1} F x, then y += C x.
template <typename T>
void schur_matvec(const HostOp<T>& op, const T* x, T* y, std::vector<T>& u, std::vector<T>& b) {
  const HostLevel<T>& L = op.levels[0];
  const int nl = op.nloc(0), m = op.nrem(0);
  u.assign(nl, T(0));
  b.assign(nl, T(0));
  for (int i = 0; i < nl; ++i) {
    T s = T(0);
    for (int k = L.F.ia[i]; k < L.F.ia[i + 1]; ++k) s += L.F.a[k] * x[L.F.ja[k]];
    u[i] = s;
  }
  solve_b(op, 0, u.data(), b.data());
  for (int i = 0; i < m; ++i) {
    T e = T(0), c = T(0);
    for (int k = L.E.ia[i]; k < L.E.ia[i + 1]; ++k) e += L.E.a[k] * b[L.E.ja[k]];
    for (int k = L.C.ia[i]; k < L.C.ia[i + 1]; ++k) c += L.C.a[k] * x[L.C.ja[k]];
    y[i] = -e + c;
  }
}

// Inner FGMRES on the level-0 Schur complement with x0 = 0 (the solve phase
// of SolveLevel with inner iterations, gemslr.cpp:3556-3563; FlexGmresClass::
// Solve, fgmres.hpp:263-558, with Krylov dimension = iteration limit =
// schur_maxits and relative tolerance schur_tol, MGS as PargemslrMgs).
template <typename T>
void schur_inner_solve(const HostOp<T>& op, const T* b, T* x, HostApplyWork<T>& w) {
  const int m = op.nrem(0);
  const int K = std::max(1, op.prm.schur_maxits);
  const double eps = 2.220446049250313e-16;
  for (int i = 0; i < m; ++i) x[i] = T(0);
  double normb = 0.0;
  {
    T s = T(0);
    for (int i = 0; i < m; ++i) s += sconj(b[i]) * b[i];
    normb = std::sqrt(abs_t(s));
  }
  if (normb < eps) return;
  // x0 = 0: the residual is b
  const double tolr = op.prm.schur_tol * normb;
  std::vector<T> V((size_t)(K + 1) * m), Z((size_t)K * m), H((size_t)(K + 1) * K, T(0)), c(K), sn(K), rs(K + 1);
  std::vector<T> t, u, bb;
  double normr = normb;
  rs[0] = normr;
  for (int r = 0; r < m; ++r) V[r] = b[r] * (1.0 / normr);
  auto Hm = [&](int i, int j) -> T& { return H[(size_t)i + (size_t)j * (K + 1)]; };
  int i = 0;
  while (i < K) {
    ++i;
    T* v = &V[(size_t)(i - 1) * m];
    T* z = &Z[(size_t)(i - 1) * m];
    T* wv = &V[(size_t)i * m];
    schur_precond(op, v, z, t, w);
    schur_matvec(op, z, wv, u, bb);
    for (int j = 0; j < i; ++j) {
      const T* vj = &V[(size_t)j * m];
      T h = T(0);
      for (int r = 0; r < m; ++r) h += sconj(vj[r]) * wv[r];
      Hm(j, i - 1) = h;
      for (int r = 0; r < m; ++r) wv[r] -= h * vj[r];
    }
    T s2 = T(0);
    for (int r = 0; r < m; ++r) s2 += sconj(wv[r]) * wv[r];
    const double tn = std::sqrt(abs_t(s2));
    Hm(i, i - 1) = tn;
    for (int r = 0; r < m; ++r) wv[r] *= T(1.0) / T(tn);
    for (int j = 1; j < i; ++j) {
      const T hii = Hm(j - 1, i - 1);
      Hm(j - 1, i - 1) = sconj(c[j - 1]) * hii + sn[j - 1] * Hm(j, i - 1);
      Hm(j, i - 1) = -sn[j - 1] * hii + c[j - 1] * Hm(j, i - 1);
    }
    const T hii = Hm(i - 1, i - 1), hii1 = Hm(i, i - 1);
    const double gam = std::sqrt(abs_t(hii) * abs_t(hii) + abs_t(hii1) * abs_t(hii1));
    if (std::fabs(gam) < eps) return;  // breakdown: x stays at the initial guess
    c[i - 1] = hii / gam;
    sn[i - 1] = hii1 / gam;
    rs[i] = -sn[i - 1] * rs[i - 1];
    rs[i - 1] = sconj(c[i - 1]) * rs[i - 1];
    Hm(i - 1, i - 1) = sconj(c[i - 1]) * hii + sn[i - 1] * hii1;
    normr = abs_t(rs[i]);
    if (normr <= tolr) break;
  }
  rs[i - 1] /= Hm(i - 1, i - 1);
  for (int k = i - 2; k >= 0; --k) {
    for (int j = k + 1; j < i; ++j) rs[k] -= Hm(k, j) * rs[j];
    rs[k] /= Hm(k, k);
  }
  for (int j = 0; j < i; ++j) {
    const T* z = &Z[(size_t)j * m];
    for (int r = 0; r < m; ++r) x[r] += rs[j] * z[r];
  }
}

}  // namespace

template <typename T>
void host_solve_level(const HostOp<T>& op, int l, const T* rhs, T* x, HostApplyWork<T>& w) {
  if (l == op.nlev - 1) {
    solve_b(op, l, rhs, x);
    return;
  }
  const HostLevel<T>& L = op.levels[l];
  const int nl = op.nloc(l), m = op.nrem(l);
  T* z = w.z[l].data();
  T* zl = z + nl;
  if (op.solve_phase && op.prm.global_solve == 1) {
    // the U variant (kGemslrUSolve, gemslr.cpp:3527-3532) skips the block L
    // solve: z_l = rhs_l
    for (int i = 0; i < m; ++i) zl[i] = rhs[nl + i];
  } else {
    solve_b(op, l, rhs, z);
    // z_l = rhs_l - E z_u
    for (int i = 0; i < m; ++i) {
      T s = T(0);
      for (int k = L.E.ia[i]; k < L.E.ia[i + 1]; ++k) s += L.E.a[k] * z[L.E.ja[k]];
      zl[i] = rhs[nl + i] - s;
    }
  }
  if (l == 0 && op.solve_phase && op.prm.schur_enable) {
    // x_l = inner FGMRES on the Schur complement (gemslr.cpp:3556-3563)
    schur_inner_solve(op, zl, x + nl, w);
  } else {
    if (L.rank > 0) {
      const int k = L.rank;
      std::vector<T>& t = w.t;
      t.assign(k, T(0));
      for (int j = 0; j < k; ++j) {
        T s = T(0);
        const T* wc = &L.W[(size_t)j * m];
        for (int i = 0; i < m; ++i) s += sconj(wc[i]) * zl[i];
        t[j] = s;
      }
      for (int j = 0; j < k; ++j) {
        const T* whc = &L.WH[(size_t)j * m];
        for (int i = 0; i < m; ++i) zl[i] += whc[i] * t[j];
      }
    }
    host_solve_level(op, l + 1, zl, x + nl, w);
  }
  // z_u = rhs_u - F x_l
  for (int i = 0; i < nl; ++i) {
    T s = T(0);
    for (int k = L.F.ia[i]; k < L.F.ia[i + 1]; ++k) s += L.F.a[k] * x[nl + L.F.ja[k]];
    z[i] = rhs[i] - s;
  }
  solve_b(op, l, z, x);
}

template <typename T>
void host_apply(const HostOp<T>& op, const T* v, T* z, HostApplyWork<T>& w) {
  const int n = op.n;
  if ((int)w.z.size() != op.nlev) {
    w.z.resize(op.nlev);
    for (int l = 0; l < op.nlev; ++l) w.z[l].assign(n - op.lev_ptr[l], T(0));
    w.r0.assign(n, T(0));
    w.x0.assign(n, T(0));
  }
  for (int i = 0; i < n; ++i) w.r0[i] = v[op.perm[i]];
  host_solve_level(op, 0, w.r0.data(), w.x0.data(), w);
  for (int i = 0; i < n; ++i) z[op.perm[i]] = w.x0[i];
}

namespace {

// G x = (I - C Ct^{-1}) x + E B^{-1} F Ct^{-1} x  (GemslrClass::EBFCMatVec,
// gemslr.cpp:3737-3792), with Ct^{-1} the GeMSLR apply of the levels below.
template <typename T>
struct EbfcOp {
  const HostOp<T>& op;
  int l;
  HostApplyWork<T> w;
  std::vector<T> zl, zu, wu;
  EbfcOp(const HostOp<T>& o, int lev) : op(o), l(lev) {
    w.z.resize(op.nlev);
    for (int q = 0; q < op.nlev; ++q) w.z[q].assign(op.n - op.lev_ptr[q], T(0));
    zl.assign(op.nrem(l), T(0));
    zu.assign(op.nloc(l), T(0));
    wu.assign(op.nloc(l), T(0));
  }
  void apply(const T* x, T* y) {
    const HostLevel<T>& L = op.levels[l];
    const int m = op.nrem(l), nl = op.nloc(l);
    host_solve_level(op, l + 1, x, zl.data(), w);
    // y = x - C zl
    parallel_for((m + 4095) / 4096, [&](int c) {
      int i1 = std::min(m, (c + 1) * 4096);
      for (int i = c * 4096; i < i1; ++i) {
        T s = T(0);
        for (int k = L.C.ia[i]; k < L.C.ia[i + 1]; ++k) s += L.C.a[k] * zl[L.C.ja[k]];
        y[i] = x[i] - s;
      }
    });
    for (int i = 0; i < nl; ++i) {
      T s = T(0);
      for (int k = L.F.ia[i]; k < L.F.ia[i + 1]; ++k) s += L.F.a[k] * zl[L.F.ja[k]];
      zu[i] = s;
    }
    solve_b(op, l, zu.data(), wu.data());
    for (int i = 0; i < m; ++i) {
      T s = T(0);
      for (int k = L.E.ia[i]; k < L.E.ia[i + 1]; ++k) s += L.E.a[k] * wu[L.E.ja[k]];
      y[i] += s;
    }
  }
};

template <typename T>
T vdot(int n, const T* x, const T* y) {  // x^H y
  T s = T(0);
  for (int i = 0; i < n; ++i) s += sconj(x[i]) * y[i];
  return s;
}

// Orthonormal basis Y (m x p) of the invariant subspace of Hm for the
// eigenvalues flagged in sel (T, Q: Schur form of Hm).  For real data the
// selection must be closed under conjugation; a real basis is built from the
// real and imaginary parts of the leading Schur vectors.
template <typename T>
int invariant_basis(int m, std::vector<dense::cplx> T_, std::vector<dense::cplx> Q, const std::vector<char>& sel,
                    std::vector<T>& Y) {
  int p = 0;
  for (char c : sel) p += c ? 1 : 0;
  if (p == 0) { Y.clear(); return 0; }
  dense::schur_reorder(m, T_, Q, sel);
  if constexpr (is_complex<T>::value) {
    Y.assign((size_t)m * p, T(0));
    for (int j = 0; j < p; ++j)
      for (int i = 0; i < m; ++i) Y[i + (size_t)j * m] = Q[i + (size_t)j * m];
    return p;
  } else {
    std::vector<double> M((size_t)m * 2 * p);
    for (int j = 0; j < p; ++j)
      for (int i = 0; i < m; ++i) {
        M[i + (size_t)j * m] = Q[i + (size_t)j * m].real();
        M[i + (size_t)(p + j) * m] = Q[i + (size_t)j * m].imag();
      }
    int r = dense::orthonormalize(m, 2 * p, M, 1e-8);
    if (r > p) r = p;
    M.resize((size_t)m * r);
    Y.swap(M);
    return r;
  }
}

// Chooses up to `budget` eigenvalue indices among `cand` (already ordered by
// preference), keeping conjugate pairs together for real data.
template <typename T>
void pick(const std::vector<dense::cplx>& lam, const std::vector<int>& cand, int budget, std::vector<char>& sel) {
  const int m = (int)lam.size();
  for (int idx : cand) {
    if (sel[idx]) continue;
    int cnt = 0;
    for (char c : sel) cnt += c ? 1 : 0;
    if (cnt >= budget) break;
    bool cplx_pair = !is_complex<T>::value && std::abs(lam[idx].imag()) > 1e-12 * std::max(1.0, std::abs(lam[idx]));
    if (!cplx_pair) { sel[idx] = 1; continue; }
    int partner = -1;
    double best = 1e300;
    for (int j = 0; j < m; ++j) {
      if (j == idx || sel[j]) continue;
      double d = std::abs(lam[j] - std::conj(lam[idx]));
      if (d < best) { best = d; partner = j; }
    }
    if (partner < 0 || cnt + 2 > budget) continue;
    sel[idx] = 1;
    sel[partner] = 1;
  }
}

// The m-length vector work of the low-rank build on the host: the basis in
// host memory, G applied by EbfcOp.
template <typename T>
struct HostLrBackend final : LrBackend<T> {
  HostOp<T>& op;
  int l = 0, m = 0;
  std::unique_ptr<EbfcOp<T>> G;
  std::vector<T> V;
  explicit HostLrBackend(HostOp<T>& o) : op(o) {}
  void begin(int lev, int mm, int ncols) override {
    l = lev;
    m = mm;
    G.reset(new EbfcOp<T>(op, l));
    V.assign((size_t)m * ncols, T(0));
  }
  T* col(int j) { return &V[(size_t)j * m]; }
  void set_col(int j, const T* x) override { std::copy(x, x + m, col(j)); }
  void apply(int j) override { G->apply(col(j), col(j + 1)); }
  void dots(int j, T* h) override {
    const T* w = col(j + 1);
    parallel_for(j + 1, [&](int q) { h[q] = vdot(m, col(q), w); });
  }
  void update(int j, const T* h) override {
    T* w = col(j + 1);
    for (int q = 0; q <= j; ++q) {
      const T* vq = col(q);
      T c = h[q];
      for (int i = 0; i < m; ++i) w[i] -= c * vq[i];
    }
  }
  double norm(int j) override { return std::sqrt(sabs(vdot(m, col(j), col(j)))); }
  void divide(int j, double d) override {
    T* w = col(j);
    for (int i = 0; i < m; ++i) w[i] /= d;
  }
  // X Y with X the first k columns of V (zero coefficients skipped)
  void combine(const T* X, int k, int p, const T* Y, int ldy, std::vector<T>& out) {
    out.assign((size_t)m * p, T(0));
    for (int j = 0; j < p; ++j)
      for (int q = 0; q < k; ++q) {
        T c = Y[q + (size_t)j * ldy];
        if (c == T(0)) continue;
        const T* xq = X + (size_t)q * m;
        T* dst = &out[(size_t)j * m];
        for (int i = 0; i < m; ++i) dst[i] += c * xq[i];
      }
  }
  void restart(int mcur, int p, const T* Y) override {
    std::vector<T> Vn;
    combine(V.data(), mcur, p, Y, mcur, Vn);
    std::memmove(col(p), col(mcur), sizeof(T) * m);
    std::memcpy(V.data(), Vn.data(), sizeof(T) * m * p);
  }
  void finish(int mcur, int r, const T* Y, const T* R, std::vector<T>& W, std::vector<T>& WH) override {
    combine(V.data(), mcur, r, Y, mcur, W);
    WH.assign((size_t)m * r, T(0));
    for (int j = 0; j < r; ++j)
      for (int q = 0; q < r; ++q) {
        T c = R[q + (size_t)j * r];
        const T* wq = &W[(size_t)q * m];
        T* dst = &WH[(size_t)j * m];
        for (int i = 0; i < m; ++i) dst[i] += wq[i] * c;
      }
  }
};

}  // namespace

// Thick-restart Arnoldi on G of level l (GemslrClass::SetupLowRankSubspace /
// thick restart, gemslr.cpp:2782-3257), the m-length work done by `be`.
template <typename T>
int lowrank_level_with(HostOp<T>& op, int l, LrBackend<T>& be) {
  HostLevel<T>& L = op.levels[l];
  const GemslrParams& P = op.prm;
  const int m = op.nrem(l);
  L.rank = 0;
  L.W.clear();
  L.WH.clear();
  L.Hk.clear();
  int rank = (l == 0) ? P.lr_rank1 : P.lr_rank2;
  if (rank <= 0 || m <= 0) return 0;
  double lr_fact = (l == 0) ? P.lr_rank_factor1 : P.lr_rank_factor2;
  double ar_fact = (l == 0) ? P.lr_arnoldi_factor1 : P.lr_arnoldi_factor2;
  double tol_eig = (l == 0) ? P.lr_tol_eig1 : P.lr_tol_eig2;
  // the reference reads lr_maxits1 on every level (gemslr.cpp:2836-2845)
  int maxits = P.lr_maxits1;
  if (maxits <= 0) return 0;
  tol_eig = std::max(tol_eig, 2.220446049250313e-16);
  int rank2 = std::max((int)(rank * lr_fact), rank);
  int lr_m = std::max((int)(rank2 * ar_fact), rank2);
  int maxsteps = std::min((int)(lr_m + lr_m * P.lr_tr_factor), m);
  rank = std::min(rank, maxsteps);
  rank2 = std::min(rank2, maxsteps);
  if (maxsteps <= 0) return 0;
  const double tol_orth = std::max(P.lr_tol_orth, 2.220446049250313e-16);

  be.begin(l, m, maxsteps + 1);
  std::vector<T> H((size_t)(maxsteps + 1) * maxsteps, T(0));
  const int ldh = maxsteps + 1;
  auto Hat = [&](int i, int j) -> T& { return H[i + (size_t)j * ldh]; };
  // starting vector
  {
    std::vector<T> v0(m);
    uint64_t s = P.seed ^ (0xA5A5ULL + 977ULL * (uint64_t)l);
    for (int i = 0; i < m; ++i) v0[i] = P.lr_rand_init ? T(unif01(s)) : T(1.0);
    double nv = std::sqrt(sabs(vdot(m, v0.data(), v0.data())));
    for (int i = 0; i < m; ++i) v0[i] /= nv;
    be.set_col(0, v0.data());
  }
  std::vector<T> h(maxsteps + 1), h2(maxsteps + 1);
  // CGS2: two rounds of dots and updates per step
  auto arnoldi = [&](int j0, int j1) -> int {
    for (int j = j0; j < j1; ++j) {
      be.apply(j);
      be.dots(j, h.data());
      be.update(j, h.data());
      be.dots(j, h2.data());
      be.update(j, h2.data());
      for (int q = 0; q <= j; ++q) Hat(q, j) = h[q] + h2[q];
      double beta = be.norm(j + 1);
      Hat(j + 1, j) = T(beta);
      if (beta < tol_orth) return j + 1;
      be.divide(j + 1, beta);
    }
    return j1;
  };

  int trlen = 0, mcur = 0;
  std::vector<dense::cplx> Tm, Qm, Ye;
  std::vector<dense::cplx> lam;
  std::vector<double> res;
  T h_last = T(0);
  bool breakdown = false;
  for (int its = 1; its <= maxits; ++its) {
    int mstepsi = std::min(lr_m + trlen, maxsteps);
    if (mstepsi <= trlen) mstepsi = std::min(trlen + 1, maxsteps);
    mcur = arnoldi(trlen, mstepsi);
    h_last = Hat(mcur, mcur - 1);
    if (sabs(h_last) < tol_orth) { h_last = T(0); breakdown = true; }
    // Schur form of Hm
    Tm.assign((size_t)mcur * mcur, 0.0);
    for (int j = 0; j < mcur; ++j)
      for (int i = 0; i < mcur; ++i) Tm[i + (size_t)j * mcur] = dense::cplx(Hat(i, j));
    std::vector<dense::cplx> Hm = Tm;
    if (dense::schur(mcur, Tm, Qm) != 0) break;
    dense::tri_eigvecs(mcur, Tm, Ye);
    lam.resize(mcur);
    res.resize(mcur);
    for (int i = 0; i < mcur; ++i) {
      lam[i] = Tm[i + (size_t)i * mcur];
      dense::cplx xl = 0.0;
      for (int j = 0; j < mcur; ++j) xl += Qm[mcur - 1 + (size_t)j * mcur] * Ye[j + (size_t)i * mcur];
      res[i] = sabs(h_last) * std::abs(xl);
    }
    std::vector<int> order(mcur);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return std::abs(lam[a]) > std::abs(lam[b]); });
    std::vector<int> cov, icov;
    for (int i : order) (res[i] < tol_eig ? cov : icov).push_back(i);
    if ((int)cov.size() >= rank2 || its == maxits || breakdown || (int)icov.size() == 0) break;
    // thick restart: converged + a fraction of the unconverged wanted ones
    int npick = std::max((int)(icov.size() * P.lr_tr_factor), 1);
    npick = std::min<int>(npick, (int)icov.size());
    std::vector<char> sel(mcur, 0);
    std::vector<int> cand(cov);
    cand.insert(cand.end(), icov.begin(), icov.begin() + npick);
    pick<T>(lam, cand, std::min((int)cov.size() + npick, mcur - 1), sel);
    std::vector<T> Y;
    int p = invariant_basis<T>(mcur, Tm, Qm, sel, Y);
    if (p <= 0 || p >= mcur) break;
    // V(:,0:p) = Vm Y ; V(:,p) = v_{m+1}
    be.restart(mcur, p, Y.data());
    // H = [Y^H Hm Y ; h_last * Y(m-1,:)]
    std::vector<T> HY((size_t)mcur * p, T(0));
    for (int j = 0; j < p; ++j)
      for (int q = 0; q < mcur; ++q) {
        T c = Y[q + (size_t)j * mcur];
        for (int i = 0; i < mcur; ++i) HY[i + (size_t)j * mcur] += from_cplx<T>(Hm[i + (size_t)q * mcur]) * c;
      }
    std::fill(H.begin(), H.end(), T(0));
    for (int j = 0; j < p; ++j) {
      for (int i = 0; i < p; ++i) {
        T s = T(0);
        for (int q = 0; q < mcur; ++q) s += sconj(Y[q + (size_t)i * mcur]) * HY[q + (size_t)j * mcur];
        Hat(i, j) = s;
      }
      Hat(p, j) = h_last * Y[mcur - 1 + (size_t)j * mcur];
    }
    trlen = p;
  }
  if (mcur <= 0 || lam.empty()) return 0;
  // final: converged eigenvalues, largest magnitude first, at most `rank`
  std::vector<int> order(mcur);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return std::abs(lam[a]) > std::abs(lam[b]); });
  std::vector<int> cov;
  for (int i : order)
    if (res[i] < tol_eig) cov.push_back(i);
  std::vector<char> sel(mcur, 0);
  pick<T>(lam, cov, rank, sel);
  std::vector<T> Y;
  // Tm/Qm hold the Schur form of the last Hm
  std::vector<dense::cplx> Hm((size_t)mcur * mcur);
  for (int j = 0; j < mcur; ++j)
    for (int i = 0; i < mcur; ++i) Hm[i + (size_t)j * mcur] = dense::cplx(Hat(i, j));
  int r = invariant_basis<T>(mcur, Tm, Qm, sel, Y);
  if (r <= 0) return 0;
  // W = Vm Y ; Tr = Y^H Hm Y ; Hk = (I - Tr)^{-1} - I ; WH = W Hk
  std::vector<T> R((size_t)r * r, T(0));
  for (int j = 0; j < r; ++j)
    for (int i = 0; i < r; ++i) {
      T s = T(0);
      for (int q = 0; q < mcur; ++q) {
        T hy = T(0);
        for (int t = 0; t < mcur; ++t) hy += from_cplx<T>(Hm[q + (size_t)t * mcur]) * Y[t + (size_t)j * mcur];
        s += sconj(Y[q + (size_t)i * mcur]) * hy;
      }
      R[i + (size_t)j * r] = (i == j ? T(1) : T(0)) - s;
    }
  if (dense::invert(r, R) != 0) return 0;
  for (int i = 0; i < r; ++i) R[i + (size_t)i * r] -= T(1);
  L.Hk = R;
  be.finish(mcur, r, Y.data(), R.data(), L.W, L.WH);
  L.rank = r;
  return r;
}

namespace {

template <typename T>
int lowrank_level(HostOp<T>& op, int l) {
  HostLrBackend<T> be(op);
  return lowrank_level_with(op, l, be);
}

}  // namespace

// Concatenated factors of one level (block-diagonal, level-local numbering)
// from its per-block ILUT factors.
template <typename T>
void build_level_cat(HostOp<T>& op, int l) {
  HostLevel<T>& L = op.levels[l];
  const int nl = op.nloc(l), s = op.lev_ptr[l];
  for (int pass = 0; pass < 2; ++pass) {
    Csr<T>& M = pass == 0 ? L.Lcat : L.Ucat;
    M = Csr<T>();
    M.nrows = M.ncols = nl;
    M.ia.assign(nl + 1, 0);
    for (int b = 0; b < L.ncomp; ++b) {
      const Csr<T>& f = pass == 0 ? L.ilu[b].L : L.ilu[b].U;
      const int off = L.dom_ptr[b] - s;
      for (int i = 0; i < f.nrows; ++i) {
        for (int k = f.ia[i]; k < f.ia[i + 1]; ++k) { M.ja.push_back(f.ja[k] + off); M.a.push_back(f.a[k]); }
        M.ia[off + i + 1] = (int)M.ja.size();
      }
    }
  }
  L.Dcat.resize(nl);
  for (int b = 0; b < L.ncomp; ++b) {
    const int off = L.dom_ptr[b] - s;
    std::copy(L.ilu[b].Dinv.begin(), L.ilu[b].Dinv.end(), L.Dcat.begin() + off);
  }
}

namespace {

int lowrank_device(HostOp<double>& op, int device, std::string* msg) {
  if (!g_lowrank_device) {
    if (msg) *msg = "no device low-rank build in this library";
    return 102;
  }
  return g_lowrank_device(op, device, msg);
}
int lowrank_device(HostOp<zdouble>& op, int device, std::string* msg) {
  if (!g_lowrank_device_z) {
    if (msg) *msg = "no device low-rank build in this library";
    return 102;
  }
  return g_lowrank_device_z(op, device, msg);
}

}  // namespace

template <typename T>
int gemslr_setup(const Csr<T>& A, const GemslrParams& prm, HostOp<T>& op, std::string* msg) {
  const double t0 = now_s();
  if (A.nrows != A.ncols) {
    if (msg) *msg = "GeMSLR setup needs a square matrix";
    return 101;
  }
  if (prm.partition != 1) {
    if (msg) *msg = "only the RKway partition (option 1) is implemented";
    return 100;
  }
  // options of the reference's parameter array this setup does not implement
  // are refused rather than replaced (gemslr.hpp:2069-2161 reads them)
  if (prm.bsolve_top != 0 || prm.bsolve_lower != 0) {
    if (msg) *msg = "only ILUT block solves (B solve option 0) are implemented";
    return 100;
  }
  if (prm.global_solve != 0 && prm.global_solve != 1) {
    if (msg) *msg = "only the LU and U variants of the level solve (global solve options 0, 1) are implemented";
    return 100;
  }
  if (prm.perm_option != 0 && prm.perm_option != 1) {
    if (msg) *msg = "only the NO and RCM local orderings (0, 1) are implemented";
    return 100;
  }
  if ((prm.lr_rank1 > 0 && prm.lr_option1 != 1) || (prm.lr_rank2 > 0 && prm.lr_option2 != 1)) {
    if (msg) *msg = "only thick-restart Arnoldi low-rank corrections (option 1) are implemented";
    return 100;
  }
  op = HostOp<T>();
  op.prm = prm;
  op.n = A.nrows;
  op.A = A;
  const int n = op.n;
  if (n == 0) {
    op.nlev = 0;
    op.lev_ptr.assign(1, 0);
    return 0;
  }
  // 1. partition and level structure
  Graph g = graph_from_pattern(n, A.ia.data(), A.ja.data());
  RkwayParams rp;
  rp.nlev = std::max(1, prm.nlev);
  rp.ncomp = prm.ncomp;
  rp.minsep = prm.minsep;
  rp.kmin = prm.kmin;
  rp.kfactor = prm.kfactor;
  rp.vertexsep = prm.vertexsep;
  rp.seed = prm.seed ^ 0x9E3779B97F4A7C15ULL;
  int nlev = 0;
  std::vector<int> map, mapptr;
  try {
    rkway_levels(g, rp, nlev, map, mapptr);
  } catch (const std::exception& e) {
    if (msg) *msg = e.what();
    return 101;
  }
  const int ndom = mapptr[nlev];
  std::vector<int> domptr(ndom + 2, 0);
  for (int i = 0; i < n; ++i) domptr[map[i] + 2]++;
  for (int d = 0; d < ndom; ++d) domptr[d + 2] += domptr[d + 1];
  op.perm.assign(n, 0);
  for (int i = 0; i < n; ++i) op.perm[domptr[map[i] + 1]++] = i;
  // now domptr[d] = start of domain d
  op.nlev = nlev;
  op.lev_ptr.assign(nlev + 1, n);
  op.levels.resize(nlev);
  for (int l = 0; l < nlev; ++l) {
    op.lev_ptr[l] = domptr[mapptr[l]];
    HostLevel<T>& L = op.levels[l];
    L.ncomp = mapptr[l + 1] - mapptr[l];
    L.dom_ptr.resize(L.ncomp + 1);
    for (int c = 0; c <= L.ncomp; ++c) L.dom_ptr[c] = domptr[mapptr[l] + c];
  }
  op.lev_ptr[nlev] = n;
  // 2. local RCM per domain
  if (prm.perm_option == 1) {
    parallel_for(ndom, [&](int d) {
      int d0 = domptr[d], nd = domptr[d + 1] - d0;
      if (nd <= 1) return;
      std::vector<int> rows(op.perm.begin() + d0, op.perm.begin() + d0 + nd);
      Csr<T> B = csr_submatrix(A, rows, rows);
      std::vector<int> p = rcm_order(B);
      for (int i = 0; i < nd; ++i) op.perm[d0 + i] = rows[p[i]];
    });
  } else if (prm.perm_option != 0) {
    if (msg) *msg = "local ordering option not supported (0: none, 1: RCM)";
    return 100;
  }
  Csr<T> Apq = csr_sym_permute(A, op.perm);
  const double t1 = now_s();
  // 3. blocks, couplings and ILUT
  struct BlockRef { int l, b; };
  std::vector<BlockRef> blocks;
  for (int l = 0; l < nlev; ++l)
    for (int b = 0; b < op.levels[l].ncomp; ++b) blocks.push_back({l, b});
  for (int l = 0; l < nlev; ++l) op.levels[l].ilu.resize(op.levels[l].ncomp);
  std::vector<int> errs(blocks.size(), 0);
  parallel_for((int)blocks.size(), [&](int t) {
    int l = blocks[t].l, b = blocks[t].b;
    HostLevel<T>& L = op.levels[l];
    int d0 = L.dom_ptr[b], nd = L.dom_ptr[b + 1] - d0;
    Csr<T> B = csr_block(Apq, d0, nd, d0, nd);
    bool last = (l == nlev - 1);
    errs[t] = ilut_factor(B, last ? prm.droptol_C : prm.droptol_B, last ? prm.rownnz_C : prm.rownnz_B, L.ilu[b]);
  });
  for (int e : errs)
    if (e) {
      if (msg) *msg = "ILUT: empty row";
      return e;
    }
  for (int l = 0; l + 1 < nlev; ++l) {
    HostLevel<T>& L = op.levels[l];
    int s = op.lev_ptr[l], e = op.lev_ptr[l + 1];
    L.E = csr_block(Apq, e, n - e, s, e - s);
    L.F = csr_block(Apq, s, e - s, e, n - e);
    int rank = (l == 0) ? prm.lr_rank1 : prm.lr_rank2;
    if (rank > 0 || (l == 0 && prm.schur_enable)) L.C = csr_block(Apq, e, n - e, e, n - e);
  }
  for (int l = 0; l < nlev; ++l) build_level_cat(op, l);
  const double t2 = now_s();
  // 4. low-rank corrections, bottom level first, on the host or on a device
  if (prm.lowrank_device >= 0) {
    const int e = lowrank_device(op, prm.lowrank_device, msg);
    if (e) return e;
  } else {
    for (int l = nlev - 2; l >= 0; --l) lowrank_level(op, l);
  }
  // the level-0 block C stays for the inner Schur iteration of the solve phase
  for (int l = 0; l + 1 < nlev; ++l)
    if (!(l == 0 && prm.schur_enable)) op.levels[l].C = Csr<T>();
  op.solve_phase = true;
  const double t3 = now_s();
  op.setup_seconds[0] = t1 - t0;
  op.setup_seconds[1] = t2 - t1;
  op.setup_seconds[2] = t3 - t2;
  op.setup_seconds[3] = t3 - t0;
  return 0;
}

LowrankDeviceFn g_lowrank_device = nullptr;
LowrankDeviceFnZ g_lowrank_device_z = nullptr;

template void build_level_cat<double>(HostOp<double>&, int);
template int lowrank_level_with<double>(HostOp<double>&, int, LrBackend<double>&);
template int lowrank_level_with<zdouble>(HostOp<zdouble>&, int, LrBackend<zdouble>&);
template void build_level_cat<zdouble>(HostOp<zdouble>&, int);
template int gemslr_setup<double>(const Csr<double>&, const GemslrParams&, HostOp<double>&, std::string*);
template int gemslr_setup<zdouble>(const Csr<zdouble>&, const GemslrParams&, HostOp<zdouble>&, std::string*);
template void host_apply<double>(const HostOp<double>&, const double*, double*, HostApplyWork<double>&);
template void host_apply<zdouble>(const HostOp<zdouble>&, const zdouble*, zdouble*, HostApplyWork<zdouble>&);
template void host_solve_level<double>(const HostOp<double>&, int, const double*, double*, HostApplyWork<double>&);
template void host_solve_level<zdouble>(const HostOp<zdouble>&, int, const zdouble*, zdouble*, HostApplyWork<zdouble>&);

}  // namespace pgs
