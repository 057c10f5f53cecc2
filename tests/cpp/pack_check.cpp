// Host emulation of the device level-set solves over the packed factors.
//
// For every level of a set-up operator the block factors are packed exactly
// as the device upload packs them (steps of at most 32 rows, ring slots, slab
// width 20) and the solve is replayed step by step with the device kernels'
// data contract: x of a block lives in a ring of `window` slots, the slot of
// the row at pair position p is p & (window - 1), a step reads every slab
// entry from the ring before any of its rows is written, the L sweep reads
// the right-hand side and the U sweep reads the L sweep's output by row id.
// Run once over whole steps (128 threads) and once over half steps (one
// wave, 16 rows per pass, the writes of the first half visible to the second).
// The result is compared with a plain forward/backward substitution on each
// block's ILUT factors: a step holding two dependent rows, or a ring slot
// reused before its last reader, changes the result.
//
// usage: pack_check   (exit 0 when every level of every problem agrees)
#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "../../pargemslr_amd/csrc/host/gemslr_host.hpp"
#include "../../pargemslr_amd/csrc/host/gen.hpp"
#include "../../pargemslr_amd/csrc/host/trsv_pack.hpp"

using namespace pgs;

namespace {

constexpr int kStepRows = 32;
constexpr int kSlabWidth = 20;

template <typename T>
T value_of(double re, double im) {
  if constexpr (is_complex<T>::value) return T(re, im);
  else return re;
}

// Forward / backward substitution on one block (the reference's
// IluClass::SolveHost order): y = L^{-1} b, x_i = dinv_i (y_i - U_i x).
template <typename T>
void block_solve(const IluFactor<T>& F, const T* b, T* x) {
  const int n = F.n;
  std::vector<T> y(n);
  for (int i = 0; i < n; ++i) {
    T s = b[i];
    for (int k = F.L.ia[i]; k < F.L.ia[i + 1]; ++k) s -= F.L.a[k] * y[F.L.ja[k]];
    y[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    T s = y[i];
    for (int k = F.U.ia[i]; k < F.U.ia[i + 1]; ++k) s -= F.U.a[k] * x[F.U.ja[k]];
    x[i] = F.Dinv[i] * s;
  }
}

// Replays one block's sweeps over the given step descriptors (whole steps or
// half steps) with a ring of W slots: within a step every read comes before
// any write, and the steps run in order.
template <typename T>
void ring_solve(const TrsvPack<T>& P, const std::vector<int>& desc, int sbeg, int sU, int send, int b, const T* rhs,
                T* x, int W) {
  std::vector<T> ring(W, T(0.0));
  const int64_t e0 = P.blk_ent0[b], q0 = P.blk_ri0[b];
  for (int pass = 0; pass < 2; ++pass) {
    const bool upper = pass == 1;
    const int s1 = upper ? sU : sbeg, s2 = upper ? send : sU;
    for (int s = s1; s < s2; ++s) {
      const int ro = desc[4 * s], eo = desc[4 * s + 1], nr = desc[4 * s + 2], w = desc[4 * s + 3];
      std::vector<T> v(nr);
      std::vector<int> row(nr);
      for (int t = 0; t < nr; ++t) {  // every read of the step first
        row[t] = P.rows[q0 + ro + t];
        T acc = T(0.0);
        for (int j = 0; j < w; ++j) {
          const int64_t e = e0 + eo + (int64_t)t * w + j;
          // col holds the slot in the packed window; a shorter ring (the
          // control) takes the same position modulo its own size
          acc += P.val[e] * ring[P.col[e] & (W - 1)];
        }
        const T src = upper ? x[row[t]] : rhs[row[t]];
        v[t] = src - acc;
        if (upper) v[t] *= P.dinv[q0 + ro + t];
      }
      for (int t = 0; t < nr; ++t) {  // then its writes
        ring[(ro + t) & (W - 1)] = v[t];
        x[row[t]] = v[t];
      }
    }
  }
}

// Returns true when every level agrees.  With `short_ring` the replay uses a
// ring of half the packed size instead (a control: it must disagree on some
// level, since the packing chose the smallest power of two that holds the
// sweep window).
template <typename T>
bool check_op(const std::string& name, const HostOp<T>& op, int max_window, bool short_ring) {
  bool ok = true;
  for (int l = 0; l < op.nlev; ++l) {
    const HostLevel<T>& L = op.levels[l];
    const int s0 = op.lev_ptr[l], nl = op.nloc(l);
    TrsvPack<T> P = trsv_pack(L, s0, nl, kStepRows, max_window, kSlabWidth);
    if (P.window == 0) {
      std::printf("%s level %d: not ring-packed (window %d needed)\n", name.c_str(), l, P.window_needed);
      ok = false;
      continue;
    }
    std::vector<int> dh, sh, suh;
    half_steps(P.desc, P.blk_step, P.blk_stepU, dh, sh, suh);
    std::vector<double> g(2 * (size_t)nl);
    rng_normal(1000 + l, 2 * (int64_t)nl, g.data());
    std::vector<T> rhs(nl), ref(nl), xw(nl), xh(nl);
    for (int i = 0; i < nl; ++i) rhs[i] = value_of<T>(g[2 * i], g[2 * i + 1]);
    int max_rows = 0;
    for (size_t i = 2; i < P.desc.size(); i += 4) max_rows = std::max(max_rows, P.desc[i]);
    for (int b = 0; b < L.ncomp; ++b) {
      const int r0 = P.blk_row0[b];
      block_solve(L.ilu[b], rhs.data() + r0, ref.data() + r0);
      const int W = short_ring ? P.window / 2 : P.window;
      ring_solve(P, P.desc, P.blk_step[b], P.blk_stepU[b], P.blk_step[b + 1], b, rhs.data() + r0, xw.data() + r0, W);
      ring_solve(P, dh, sh[b], suh[b], sh[b + 1], b, rhs.data() + r0, xh.data() + r0, W);
    }
    double emax = 0.0, nmax = 0.0;
    for (int i = 0; i < nl; ++i) {
      nmax = std::max(nmax, std::abs(ref[i]));
      emax = std::max(emax, std::max(std::abs(xw[i] - ref[i]), std::abs(xh[i] - ref[i])));
    }
    const double rel = nmax > 0.0 ? emax / nmax : emax;
    const bool good = rel <= 1e-12 && max_rows <= kStepRows;
    std::printf("%s level %d: %d blocks, %zu steps, %zu half steps, window %d%s, widest step %d rows, rel err "
                "%.2e %s\n",
                name.c_str(), l, L.ncomp, P.desc.size() / 4, dh.size() / 4, P.window,
                short_ring ? " (replayed on half)" : "", max_rows, rel, good ? "ok" : "MISMATCH");
    ok = ok && good;
  }
  return ok;
}

template <typename T>
bool run(const std::string& name, const Csr<T>& A, int nlev, int ncomp, double droptol, int rownnz, int rank,
         int max_window, bool short_ring = false) {
  GemslrParams prm;
  prm.nlev = nlev;
  prm.ncomp = ncomp;
  prm.droptol_B = prm.droptol_C = droptol;
  prm.rownnz_B = prm.rownnz_C = rownnz;
  prm.lr_rank1 = prm.lr_rank2 = rank;
  prm.lr_tol_eig1 = prm.lr_tol_eig2 = 1e-3;
  prm.lr_maxits1 = prm.lr_maxits2 = 10;
  prm.schur_enable = 0;
  HostOp<T> op;
  std::string msg;
  if (gemslr_setup(A, prm, op, &msg)) {
    std::printf("%s: setup failed: %s\n", name.c_str(), msg.c_str());
    return false;
  }
  return check_op(name, op, max_window, short_ring);
}

}  // namespace

int main() {
  bool ok = true;
  // rings as large as the device's LDS allows for FP64 values (8192 slots);
  // the packing takes the smallest power of two that holds the sweep window
  const int mw = 8192;
  ok = run("laplacian 24^3", laplacian<double>(24, 24, 24, 0.0, 0.0, 0.0, 0.0), 3, 8, 1e-3, 20, 8, mw) && ok;
  ok = run("laplacian 32^3 convective", laplacian<double>(32, 32, 32, 0.0, 0.05, -0.02, 0.01), 3, 8, 1e-3, 20, 8,
           mw) && ok;
  ok = run("convdiff 20^3", convdiff_upwind(20, 20, 20, 0.3), 2, 4, 1e-2, 10, 0, mw) && ok;
  ok = run("complex laplacian 16^3", laplacian<zdouble>(16, 16, 16, zdouble(0.2, 0.02), 0.05, -0.02, 0.01), 3, 4,
           1e-3, 20, 8, mw) && ok;
  // control: a ring of half the packed size must break the replay
  const bool control = run("control 24^3", laplacian<double>(24, 24, 24, 0.0, 0.0, 0.0, 0.0), 3, 8, 1e-3, 20, 8, mw,
                           true);
  std::printf("%s\n", control ? "CONTROL NOT DETECTED" : "control detected");
  ok = ok && !control;
  std::printf("%s\n", ok ? "all levels agree" : "MISMATCH");
  return ok ? 0 : 1;
}
