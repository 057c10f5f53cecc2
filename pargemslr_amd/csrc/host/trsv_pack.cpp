#include "trsv_pack.hpp"

#include <algorithm>

namespace pgs {

namespace {

template <typename T>
void pack_sweep(const Csr<T>& M, const std::vector<T>* dinv, bool upper, int max_rows, int64_t ent0, int64_t ri0,
                TrsvPack<T>& P) {
  const int n = M.nrows;
  std::vector<int> lev(n, 0);
  int nl = 0;
  if (!upper) {
    for (int i = 0; i < n; ++i) {
      int l = 0;
      for (int k = M.ia[i]; k < M.ia[i + 1]; ++k) l = std::max(l, lev[M.ja[k]] + 1);
      lev[i] = l;
      nl = std::max(nl, l + 1);
    }
  } else {
    for (int i = n - 1; i >= 0; --i) {
      int l = 0;
      for (int k = M.ia[i]; k < M.ia[i + 1]; ++k) l = std::max(l, lev[M.ja[k]] + 1);
      lev[i] = l;
      nl = std::max(nl, l + 1);
    }
  }
  // bucket rows by level (stable in row order)
  std::vector<int> cnt(nl + 1, 0);
  for (int i = 0; i < n; ++i) cnt[lev[i] + 1]++;
  for (int l = 0; l < nl; ++l) cnt[l + 1] += cnt[l];
  std::vector<int> order(n), pos(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; ++i) order[pos[lev[i]]++] = i;
  for (int l = 0; l < nl; ++l) {
    for (int c0 = cnt[l]; c0 < cnt[l + 1]; c0 += max_rows) {
      const int nr = std::min(max_rows, cnt[l + 1] - c0);
      int width = 0;
      for (int t = 0; t < nr; ++t) {
        int i = order[c0 + t];
        width = std::max(width, M.ia[i + 1] - M.ia[i]);
      }
      const int64_t row_off = (int64_t)P.rows.size() - ri0, ent_off = (int64_t)P.col.size() - ent0;
      P.desc.push_back((int)row_off);
      P.desc.push_back((int)ent_off);
      P.desc.push_back(nr);
      P.desc.push_back(width);
      P.max_width = std::max(P.max_width, width);
      const size_t base = P.col.size();
      P.col.resize(base + (size_t)width * nr, 0);
      P.val.resize(base + (size_t)width * nr, T(0));
      for (int t = 0; t < nr; ++t) {
        int i = order[c0 + t];
        P.rows.push_back(i);
        P.dinv.push_back(dinv ? (*dinv)[i] : T(0));
        int j = 0;
        for (int k = M.ia[i]; k < M.ia[i + 1]; ++k, ++j) {
          P.col[base + (size_t)j * nr + t] = M.ja[k];
          P.val[base + (size_t)j * nr + t] = M.a[k];
        }
        P.nnz_real += M.ia[i + 1] - M.ia[i];
      }
    }
  }
}

}  // namespace

template <typename T>
TrsvPack<T> trsv_pack(const HostLevel<T>& L, int lev_start, int nloc, int max_rows) {
  TrsvPack<T> P;
  P.nblocks = L.ncomp;
  P.nloc = nloc;
  P.blk_row0.resize(L.ncomp + 1);
  P.blk_step.resize(L.ncomp + 1);
  P.blk_stepU.resize(L.ncomp);
  P.blk_ent0.resize(L.ncomp + 1);
  P.blk_ri0.resize(L.ncomp + 1);
  for (int b = 0; b < L.ncomp; ++b) {
    P.blk_row0[b] = L.dom_ptr[b] - lev_start;
    const IluFactor<T>& F = L.ilu[b];
    P.max_block_rows = std::max(P.max_block_rows, F.n);
    const int64_t ent0 = (int64_t)P.col.size(), ri0 = (int64_t)P.rows.size();
    P.blk_ent0[b] = ent0;
    P.blk_ri0[b] = ri0;
    P.blk_step[b] = (int)P.desc.size() / 4;
    pack_sweep<T>(F.L, nullptr, false, max_rows, ent0, ri0, P);
    P.blk_stepU[b] = (int)P.desc.size() / 4;
    pack_sweep<T>(F.U, &F.Dinv, true, max_rows, ent0, ri0, P);
    P.max_steps = std::max(P.max_steps, (int)P.desc.size() / 4 - P.blk_step[b]);
  }
  P.blk_row0[L.ncomp] = nloc;
  P.blk_step[L.ncomp] = (int)P.desc.size() / 4;
  P.blk_ent0[L.ncomp] = (int64_t)P.col.size();
  P.blk_ri0[L.ncomp] = (int64_t)P.rows.size();
  return P;
}

template TrsvPack<double> trsv_pack<double>(const HostLevel<double>&, int, int, int);
template TrsvPack<zdouble> trsv_pack<zdouble>(const HostLevel<zdouble>&, int, int, int);

}  // namespace pgs
