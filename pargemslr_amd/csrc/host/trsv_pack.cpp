#include "trsv_pack.hpp"

#include <algorithm>

namespace pgs {

namespace {

// Rows of one triangular factor in sweep order, cut into steps.
struct SweepOrder {
  std::vector<int> order;  // rows in sweep order
  std::vector<int> step;   // step boundaries into order (nsteps + 1)
  std::vector<int> pos;    // pos[row] = index in order
  int window = 0;
};

// Forward topological order: L rows increasing (dependencies j < i), U rows
// decreasing (dependencies j > i).
inline int topo_row(int s, int n, bool upper) { return upper ? n - 1 - s : s; }

template <typename T>
int levels_asap(const Csr<T>& M, bool upper, std::vector<int>& lev) {
  const int n = M.nrows;
  lev.assign(n, 0);
  int nl = 0;
  for (int s = 0; s < n; ++s) {
    const int i = topo_row(s, n, upper);
    int l = 0;
    for (int k = M.ia[i]; k < M.ia[i + 1]; ++k) l = std::max(l, lev[M.ja[k]] + 1);
    lev[i] = l;
    nl = std::max(nl, l + 1);
  }
  return nl;
}

template <typename T>
void levels_alap(const Csr<T>& M, bool upper, int nl, std::vector<int>& lev) {
  const int n = M.nrows;
  lev.assign(n, std::max(nl - 1, 0));
  for (int s = n - 1; s >= 0; --s) {
    const int i = topo_row(s, n, upper);
    for (int k = M.ia[i]; k < M.ia[i + 1]; ++k) lev[M.ja[k]] = std::min(lev[M.ja[k]], lev[i] - 1);
  }
}

template <typename T>
SweepOrder make_order(const Csr<T>& M, const std::vector<int>& lev, int nl, int max_rows) {
  const int n = M.nrows;
  SweepOrder o;
  std::vector<int> cnt(nl + 1, 0);
  for (int i = 0; i < n; ++i) cnt[lev[i] + 1]++;
  for (int l = 0; l < nl; ++l) cnt[l + 1] += cnt[l];
  o.order.resize(n);
  o.pos.resize(n);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; ++i) o.order[fill[lev[i]]++] = i;  // stable in row order
  for (int k = 0; k < n; ++k) o.pos[o.order[k]] = k;
  o.step.push_back(0);
  for (int l = 0; l < nl; ++l)
    for (int c0 = cnt[l]; c0 < cnt[l + 1]; c0 += max_rows) o.step.push_back(std::min(cnt[l + 1], c0 + max_rows));
  // window: largest (end of a step) - (position of a row it reads)
  for (size_t s = 0; s + 1 < o.step.size(); ++s) {
    const int b = o.step[s + 1];
    for (int k = o.step[s]; k < b; ++k) {
      const int i = o.order[k];
      for (int e = M.ia[i]; e < M.ia[i + 1]; ++e) o.window = std::max(o.window, b - o.pos[M.ja[e]]);
    }
  }
  return o;
}

// The schedule of the two with the smaller window (as soon as possible on a tie).
template <typename T>
SweepOrder best_order(const Csr<T>& M, bool upper, int max_rows) {
  std::vector<int> lev;
  const int nl = levels_asap(M, upper, lev);
  SweepOrder a = make_order(M, lev, nl, max_rows);
  if (a.window == 0) return a;
  levels_alap(M, upper, nl, lev);
  SweepOrder b = make_order(M, lev, nl, max_rows);
  return b.window < a.window ? b : a;
}

// Slab width of each step of a sweep (fixed_width when the rows are
// ring-packed, else the longest row of the step).
template <typename T>
std::vector<int> step_widths(const Csr<T>& M, const SweepOrder& o, int fixed_width) {
  std::vector<int> w(o.step.size() - 1, fixed_width);
  if (fixed_width) return w;
  for (size_t s = 0; s + 1 < o.step.size(); ++s)
    for (int k = o.step[s]; k < o.step[s + 1]; ++k) {
      const int i = o.order[k];
      w[s] = std::max(w[s], M.ia[i + 1] - M.ia[i]);
    }
  return w;
}

// Writes the steps of one sweep at the given offsets of P's arrays (steps
// from step0, entries from ent, rows from row); entry and row offsets in the
// descriptors are relative to the block's first entry and row (ent0, ri0).
template <typename T>
void emit_sweep(const Csr<T>& M, const std::vector<T>* dinv, const SweepOrder& o, const std::vector<int>& width,
                int base, int wmask, int64_t ent0, int64_t ri0, int64_t step0, int64_t ent, int64_t row,
                TrsvPack<T>& P) {
  for (size_t s = 0; s + 1 < o.step.size(); ++s) {
    const int k0 = o.step[s], nr = o.step[s + 1] - k0, w = width[s];
    int* d = &P.desc[(size_t)(step0 + (int64_t)s) * 4];
    d[0] = (int)(row - ri0);
    d[1] = (int)(ent - ent0);
    d[2] = nr;
    d[3] = w;
    for (int t = 0; t < nr; ++t) {
      const int i = o.order[k0 + t];
      P.rows[row + t] = i;
      P.dinv[row + t] = dinv ? (*dinv)[i] : T(0);
      int j = 0;
      for (int k = M.ia[i]; k < M.ia[i + 1]; ++k, ++j) {
        const int c = M.ja[k];
        P.col[ent + (int64_t)t * w + j] = wmask ? ((base + o.pos[c]) & wmask) : c;
        P.val[ent + (int64_t)t * w + j] = M.a[k];
      }
    }
    row += nr;
    ent += (int64_t)w * nr;
  }
}

}  // namespace

template <typename T>
TrsvPack<T> trsv_pack(const HostLevel<T>& L, int lev_start, int nloc, int max_rows, int max_window,
                      int ring_width) {
  TrsvPack<T> P;
  P.nblocks = L.ncomp;
  P.nloc = nloc;
  std::vector<SweepOrder> oL(L.ncomp), oU(L.ncomp);
  // the blocks' schedules are independent: one task per block and factor
  parallel_for(2 * L.ncomp, [&](int t) {
    const IluFactor<T>& F = L.ilu[t / 2];
    if (t % 2 == 0) oL[t / 2] = best_order(F.L, false, max_rows);
    else oU[t / 2] = best_order(F.U, true, max_rows);
  });
  for (int b = 0; b < L.ncomp; ++b)
    P.window_needed = std::max(P.window_needed, std::max(oL[b].window, oU[b].window));
  int W = 1;
  while (W < P.window_needed) W <<= 1;
  int row_len = 0;
  for (const IluFactor<T>& F : L.ilu)
    for (const Csr<T>* M : {&F.L, &F.U})
      for (int i = 0; i < M->nrows; ++i) row_len = std::max(row_len, M->ia[i + 1] - M->ia[i]);
  P.window = (max_window > 0 && W <= max_window && row_len <= ring_width) ? W : 0;
  const int fixed_width = P.window ? ring_width : 0;
  const int wmask = P.window ? P.window - 1 : 0;
  P.blk_row0.resize(L.ncomp + 1);
  P.blk_step.resize(L.ncomp + 1);
  P.blk_stepU.resize(L.ncomp);
  P.blk_ent0.resize(L.ncomp + 1);
  P.blk_ri0.resize(L.ncomp + 1);
  // sizes of every block, then one allocation and the blocks emitted in
  // parallel at their offsets
  std::vector<std::vector<int>> wL(L.ncomp), wU(L.ncomp);
  std::vector<int64_t> nst(L.ncomp), nstL(L.ncomp), nent(L.ncomp), nrow(L.ncomp), nreal(L.ncomp);
  std::vector<int> maxw(L.ncomp, 0);
  parallel_for(L.ncomp, [&](int b) {
    const IluFactor<T>& F = L.ilu[b];
    wL[b] = step_widths(F.L, oL[b], fixed_width);
    wU[b] = step_widths(F.U, oU[b], fixed_width);
    int64_t e = 0;
    for (size_t s = 0; s < wL[b].size(); ++s) e += (int64_t)wL[b][s] * (oL[b].step[s + 1] - oL[b].step[s]);
    for (size_t s = 0; s < wU[b].size(); ++s) e += (int64_t)wU[b][s] * (oU[b].step[s + 1] - oU[b].step[s]);
    for (int w : wL[b]) maxw[b] = std::max(maxw[b], w);
    for (int w : wU[b]) maxw[b] = std::max(maxw[b], w);
    nstL[b] = (int64_t)wL[b].size();
    nst[b] = nstL[b] + (int64_t)wU[b].size();
    nent[b] = e;
    nrow[b] = (int64_t)F.L.nrows + F.U.nrows;
    nreal[b] = (int64_t)F.L.nnz() + F.U.nnz();
  });
  int64_t st = 0, en = 0, rw = 0;
  for (int b = 0; b < L.ncomp; ++b) {
    P.blk_row0[b] = L.dom_ptr[b] - lev_start;
    P.max_block_rows = std::max(P.max_block_rows, L.ilu[b].n);
    P.blk_ent0[b] = en;
    P.blk_ri0[b] = rw;
    P.blk_step[b] = (int)st;
    P.blk_stepU[b] = (int)(st + nstL[b]);
    P.max_steps = std::max(P.max_steps, (int)nst[b]);
    P.max_width = std::max(P.max_width, maxw[b]);
    P.nnz_real += nreal[b];
    st += nst[b];
    en += nent[b];
    rw += nrow[b];
  }
  P.blk_row0[L.ncomp] = nloc;
  P.blk_step[L.ncomp] = (int)st;
  P.blk_ent0[L.ncomp] = en;
  P.blk_ri0[L.ncomp] = rw;
  P.desc.assign((size_t)st * 4, 0);
  P.col.assign((size_t)en, 0);
  P.val.assign((size_t)en, T(0));
  P.rows.assign((size_t)rw, 0);
  P.dinv.assign((size_t)rw, T(0));
  parallel_for(L.ncomp, [&](int b) {
    const IluFactor<T>& F = L.ilu[b];
    const int64_t e0 = P.blk_ent0[b], r0 = P.blk_ri0[b];
    int64_t eU = e0;
    for (size_t s = 0; s < wL[b].size(); ++s) eU += (int64_t)wL[b][s] * (oL[b].step[s + 1] - oL[b].step[s]);
    emit_sweep<T>(F.L, nullptr, oL[b], wL[b], 0, wmask, e0, r0, P.blk_step[b], e0, r0, P);
    emit_sweep<T>(F.U, &F.Dinv, oU[b], wU[b], F.n, wmask, e0, r0, P.blk_stepU[b], eU, r0 + F.L.nrows, P);
    oL[b] = SweepOrder();
    oU[b] = SweepOrder();
  });
  return P;
}

template TrsvPack<double> trsv_pack<double>(const HostLevel<double>&, int, int, int, int, int);
template TrsvPack<zdouble> trsv_pack<zdouble>(const HostLevel<zdouble>&, int, int, int, int, int);

void half_steps(const std::vector<int>& desc, const std::vector<int>& blk_step, const std::vector<int>& blk_stepU,
                std::vector<int>& desc_h, std::vector<int>& step_h, std::vector<int>& stepU_h) {
  const int nb = (int)blk_step.size() - 1;
  desc_h.clear();
  step_h.assign(nb + 1, 0);
  stepU_h.assign(nb, 0);
  auto push = [&](int x, int y, int z, int w) {
    desc_h.push_back(x);
    desc_h.push_back(y);
    desc_h.push_back(z);
    desc_h.push_back(w);
  };
  for (int b = 0; b < nb; ++b) {
    step_h[b] = (int)(desc_h.size() / 4);
    for (int s = blk_step[b]; s < blk_step[b + 1]; ++s) {
      if (s == blk_stepU[b]) stepU_h[b] = (int)(desc_h.size() / 4);
      const int x = desc[4 * s], y = desc[4 * s + 1], z = desc[4 * s + 2], w = desc[4 * s + 3];
      push(x, y, std::min(z, 16), w);
      if (z > 16) push(x + 16, y + 16 * w, z - 16, w);
    }
    if (blk_stepU[b] == blk_step[b + 1]) stepU_h[b] = (int)(desc_h.size() / 4);
  }
  step_h[nb] = (int)(desc_h.size() / 4);
}

}  // namespace pgs
