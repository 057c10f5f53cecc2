#include "trsv_pack.hpp"

#include <algorithm>

namespace pgs {

namespace {

template <typename T>
void pack_sweep(const Csr<T>& M, const std::vector<T>* dinv, bool upper, TrsvPack<T>& P) {
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
    const int r0 = cnt[l], nr = cnt[l + 1] - r0;
    int width = 0;
    for (int t = 0; t < nr; ++t) {
      int i = order[r0 + t];
      width = std::max(width, M.ia[i + 1] - M.ia[i]);
    }
    const int row_off = (int)P.rows.size(), ent_off = (int)P.col.size();
    P.desc.push_back(row_off);
    P.desc.push_back(ent_off);
    P.desc.push_back(nr);
    P.desc.push_back(width);
    P.col.resize(ent_off + (size_t)width * nr, 0);
    P.val.resize(ent_off + (size_t)width * nr, T(0));
    for (int t = 0; t < nr; ++t) {
      int i = order[r0 + t];
      P.rows.push_back(i);
      P.dinv.push_back(dinv ? (*dinv)[i] : T(0));
      int j = 0;
      for (int k = M.ia[i]; k < M.ia[i + 1]; ++k, ++j) {
        P.col[ent_off + (size_t)j * nr + t] = M.ja[k];
        P.val[ent_off + (size_t)j * nr + t] = M.a[k];
      }
      P.nnz_real += M.ia[i + 1] - M.ia[i];
    }
  }
}

}  // namespace

template <typename T>
TrsvPack<T> trsv_pack(const HostLevel<T>& L, int lev_start, int nloc) {
  TrsvPack<T> P;
  P.nblocks = L.ncomp;
  P.nloc = nloc;
  P.blk_row0.resize(L.ncomp + 1);
  P.blk_setL.resize(L.ncomp + 1);
  P.blk_setU.resize(L.ncomp + 1);
  for (int b = 0; b < L.ncomp; ++b) {
    P.blk_row0[b] = L.dom_ptr[b] - lev_start;
    const IluFactor<T>& F = L.ilu[b];
    P.max_block_rows = std::max(P.max_block_rows, F.n);
    P.blk_setL[b] = (int)P.desc.size() / 4;
    pack_sweep<T>(F.L, nullptr, false, P);
    P.blk_setU[b] = (int)P.desc.size() / 4;
    int nsL = P.blk_setU[b] - P.blk_setL[b];
    pack_sweep(F.U, &F.Dinv, true, P);
    int nsU = (int)P.desc.size() / 4 - P.blk_setU[b];
    P.max_sets = std::max(P.max_sets, std::max(nsL, nsU));
  }
  P.blk_row0[L.ncomp] = nloc;
  // blk_setL[b+1] is not the end of block b's L sets: U sets follow them.  The
  // kernel uses [blk_setL[b], blk_setU[b]) for L and [blk_setU[b], blk_setL[b+1]) for U.
  P.blk_setL[L.ncomp] = (int)P.desc.size() / 4;
  P.blk_setU[L.ncomp] = (int)P.desc.size() / 4;
  return P;
}

template TrsvPack<double> trsv_pack<double>(const HostLevel<double>&, int, int);
template TrsvPack<zdouble> trsv_pack<zdouble>(const HostLevel<zdouble>&, int, int);

}  // namespace pgs
