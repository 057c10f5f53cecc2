#include "rcm_ilut.hpp"

#include <algorithm>
#include <queue>

namespace pgs {

namespace {

struct Pattern {
  int n = 0;
  std::vector<int> xadj, adj;
};

template <typename T>
Pattern sym_pattern(const Csr<T>& A) {
  Pattern g;
  g.n = A.nrows;
  std::vector<int> cnt(g.n + 1, 0);
  for (int i = 0; i < g.n; ++i)
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k) {
      int j = A.ja[k];
      if (j == i) continue;
      cnt[i + 1]++;
      cnt[j + 1]++;
    }
  for (int i = 0; i < g.n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> tmp(cnt[g.n]), pos(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < g.n; ++i)
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k) {
      int j = A.ja[k];
      if (j == i) continue;
      tmp[pos[i]++] = j;
      tmp[pos[j]++] = i;
    }
  g.xadj.assign(g.n + 1, 0);
  for (int i = 0; i < g.n; ++i) {
    auto b = tmp.begin() + cnt[i], e = tmp.begin() + cnt[i + 1];
    std::sort(b, e);
    auto last = std::unique(b, e);
    g.adj.insert(g.adj.end(), b, last);
    g.xadj[i + 1] = (int)g.adj.size();
  }
  return g;
}

// BFS levels from root over unnumbered vertices (numbered[v] == 0).
int bfs_levels(const Pattern& g, int root, const std::vector<char>& numbered, std::vector<int>& stamp, int tag,
               std::vector<int>& order, std::vector<int>& lev_start) {
  order.clear();
  lev_start.clear();
  order.push_back(root);
  stamp[root] = tag;
  lev_start.push_back(0);
  size_t b = 0, e = 1;
  while (b < e) {
    for (size_t h = b; h < e; ++h) {
      int v = order[h];
      for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        int u = g.adj[k];
        if (numbered[u] || stamp[u] == tag) continue;
        stamp[u] = tag;
        order.push_back(u);
      }
    }
    b = e;
    e = order.size();
    if (e > b) lev_start.push_back((int)b);
  }
  lev_start.push_back((int)order.size());
  return (int)lev_start.size() - 1;  // number of levels
}

}  // namespace

template <typename T>
std::vector<int> rcm_order(const Csr<T>& A) {
  const int n = A.nrows;
  std::vector<int> perm;
  perm.reserve(n);
  if (n == 0) return perm;
  Pattern g = sym_pattern(A);
  std::vector<char> numbered(n, 0);
  std::vector<int> stamp(n, -1), order, lev;
  int tag = 0, first = 0;
  while ((int)perm.size() < n) {
    while (numbered[first]) ++first;
    // root: minimum-degree vertex of the component containing `first`
    bfs_levels(g, first, numbered, stamp, tag++, order, lev);
    int root = first, mind = n + 1;
    for (int v : order) {
      int d = g.xadj[v + 1] - g.xadj[v];
      if (d < mind) { mind = d; root = v; }
    }
    // pseudo-peripheral vertex
    int nlev = bfs_levels(g, root, numbered, stamp, tag++, order, lev);
    for (;;) {
      int best = root, bd = n + 1;
      for (int h = lev[nlev - 1]; h < lev[nlev]; ++h) {
        int v = order[h];
        int d = g.xadj[v + 1] - g.xadj[v];
        if (d < bd) { bd = d; best = v; }
      }
      std::vector<int> order2, lev2;
      int nlev2 = bfs_levels(g, best, numbered, stamp, tag++, order2, lev2);
      root = best;
      if (nlev2 <= nlev) { order.swap(order2); lev.swap(lev2); nlev = nlev2; break; }
      order.swap(order2);
      lev.swap(lev2);
      nlev = nlev2;
    }
    // BFS numbering from root, reversed per component
    for (int v : order) numbered[v] = 1;
    perm.insert(perm.end(), order.rbegin(), order.rend());
  }
  return perm;
}

template <typename T>
int ilut_factor(const Csr<T>& A, double droptol, int lfil, IluFactor<T>& F) {
  const int n = A.nrows;
  F.n = n;
  F.L = Csr<T>();
  F.U = Csr<T>();
  F.L.nrows = F.L.ncols = n;
  F.U.nrows = F.U.ncols = n;
  F.L.ia.assign(n + 1, 0);
  F.U.ia.assign(n + 1, 0);
  F.Dinv.assign(n, T(0));
  if (lfil < 0) lfil = 0;
  // working row: dense value array + position marker
  std::vector<T> w(n, T(0));
  std::vector<int> pos(n, -1);
  std::vector<int> lcols, ucols;
  std::priority_queue<int, std::vector<int>, std::greater<int>> pending;
  std::vector<std::pair<double, int>> sel;
  for (int i = 0; i < n; ++i) {
    double tnorm = 0.0;
    lcols.clear();
    ucols.clear();
    T diag = T(0);
    int nrow = A.ia[i + 1] - A.ia[i];
    for (int k = A.ia[i]; k < A.ia[i + 1]; ++k) {
      int c = A.ja[k];
      T v = A.a[k];
      tnorm += sabs(v);
      if (c < i) {
        if (pos[c] < 0) { pos[c] = 1; lcols.push_back(c); pending.push(c); w[c] = v; }
        else w[c] += v;
      } else if (c == i) {
        diag += v;
      } else {
        if (pos[c] < 0) { pos[c] = 1; ucols.push_back(c); w[c] = v; }
        else w[c] += v;
      }
    }
    if (tnorm == 0.0) {
      for (int c : lcols) pos[c] = -1;
      for (int c : ucols) pos[c] = -1;
      while (!pending.empty()) pending.pop();
      return 103;
    }
    const double tolnorm = droptol * tnorm / (double)nrow;
    pos[i] = 1;
    w[i] = diag;
    // eliminate with previous rows in increasing column order
    while (!pending.empty()) {
      int jrow = pending.top();
      pending.pop();
      T fact = w[jrow] * F.Dinv[jrow];
      w[jrow] = fact;
      for (int k = F.U.ia[jrow]; k < F.U.ia[jrow + 1]; ++k) {
        int c = F.U.ja[k];
        T lxu = -fact * F.U.a[k];
        if (pos[c] < 0) {
          if (sabs(lxu) < tolnorm) continue;  // small fill-in dropped
          pos[c] = 1;
          w[c] = lxu;
          if (c < i) { lcols.push_back(c); pending.push(c); }
          else ucols.push_back(c);
        } else {
          w[c] += lxu;
        }
      }
    }
    // keep the lfil largest of L and of U
    auto store = [&](std::vector<int>& cols, Csr<T>& M) {
      int len = std::min<int>((int)cols.size(), lfil);
      sel.clear();
      for (int c : cols) sel.emplace_back(sabs(w[c]), c);
      if (len < (int)sel.size()) {
        std::nth_element(sel.begin(), sel.begin() + len, sel.end(),
                         [](const std::pair<double, int>& x, const std::pair<double, int>& y) {
                           return x.first > y.first || (x.first == y.first && x.second < y.second);
                         });
        sel.resize(len);
      }
      std::sort(sel.begin(), sel.end(), [](const std::pair<double, int>& x, const std::pair<double, int>& y) { return x.second < y.second; });
      for (auto& s : sel) { M.ja.push_back(s.second); M.a.push_back(w[s.second]); }
      M.ia[i + 1] = (int)M.ja.size();
    };
    store(lcols, F.L);
    store(ucols, F.U);
    T d = w[i];
    if (sabs(d) < 1e-6) d = T(1e-6);
    F.Dinv[i] = T(1.0) / d;
    for (int c : lcols) { pos[c] = -1; w[c] = T(0); }
    for (int c : ucols) { pos[c] = -1; w[c] = T(0); }
    pos[i] = -1;
    w[i] = T(0);
  }
  return 0;
}

template std::vector<int> rcm_order<double>(const Csr<double>&);
template std::vector<int> rcm_order<zdouble>(const Csr<zdouble>&);
template int ilut_factor<double>(const Csr<double>&, double, int, IluFactor<double>&);
template int ilut_factor<zdouble>(const Csr<zdouble>&, double, int, IluFactor<zdouble>&);

}  // namespace pgs
