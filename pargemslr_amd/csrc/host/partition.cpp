// Multilevel graph bisection, vertex separators, recursive k-way and the
// GeMSLR level structure.  Algorithms follow the published multilevel scheme
// (coarsen by heavy-edge matching, grow an initial bisection, project and
// refine with boundary Fiduccia-Mattheyses); the level-structure driver follows
// SetupPermutationRKwayRecursive (ParGeMSLR/SRC/matrices/matrixops.cpp:11454-11692)
// and SetupPermutationNDRecursive (matrixops.cpp:11736-11911).
#include "partition.hpp"

#include <algorithm>
#include <cstring>
#include <queue>
#include <stdexcept>

namespace pgs {

namespace {

inline uint64_t mix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline uint64_t derive_seed(uint64_t seed, uint64_t tag) {
  uint64_t s = seed ^ (tag * 0xD1B54A32D192ED03ULL);
  return mix64(s);
}

void random_perm(int n, uint64_t& rng, std::vector<int>& p) {
  p.resize(n);
  for (int i = 0; i < n; ++i) p[i] = i;
  for (int i = n - 1; i > 0; --i) {
    int j = (int)(mix64(rng) % (uint64_t)(i + 1));
    std::swap(p[i], p[j]);
  }
}

// Heavy-edge matching coarsening.  cmap[v] = coarse vertex of v.
Graph coarsen(const Graph& g, uint64_t& rng, std::vector<int>& cmap) {
  const int n = g.n;
  std::vector<int> match(n, -1), perm;
  random_perm(n, rng, perm);
  for (int v : perm) {
    if (match[v] >= 0) continue;
    int best = -1, bestw = -1;
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int u = g.adj[k];
      if (match[u] < 0 && u != v && g.ew[k] > bestw) { bestw = g.ew[k]; best = u; }
    }
    if (best >= 0) { match[v] = best; match[best] = v; }
    else match[v] = v;
  }
  cmap.assign(n, -1);
  int nc = 0;
  std::vector<int> first;  // a representative fine vertex per coarse vertex
  first.reserve(n / 2 + 1);
  for (int v : perm) {
    if (cmap[v] >= 0) continue;
    cmap[v] = nc;
    cmap[match[v]] = nc;
    first.push_back(v);
    ++nc;
  }
  Graph c;
  c.n = nc;
  c.vw.assign(nc, 0);
  c.xadj.assign(nc + 1, 0);
  c.adj.reserve(g.adj.size() / 2 + 16);
  c.ew.reserve(g.adj.size() / 2 + 16);
  std::vector<int> pos(nc, -1);
  for (int cv = 0; cv < nc; ++cv) {
    int v = first[cv];
    int members[2] = {v, match[v]};
    int nm = (match[v] == v) ? 1 : 2;
    int start = (int)c.adj.size();
    for (int m = 0; m < nm; ++m) {
      int f = members[m];
      c.vw[cv] += g.vw[f];
      for (int k = g.xadj[f]; k < g.xadj[f + 1]; ++k) {
        int cu = cmap[g.adj[k]];
        if (cu == cv) continue;
        if (pos[cu] < 0) {
          pos[cu] = (int)c.adj.size();
          c.adj.push_back(cu);
          c.ew.push_back(g.ew[k]);
        } else {
          c.ew[pos[cu]] += g.ew[k];
        }
      }
    }
    for (int k = start; k < (int)c.adj.size(); ++k) pos[c.adj[k]] = -1;
    c.xadj[cv + 1] = (int)c.adj.size();
  }
  return c;
}

int64_t edge_cut(const Graph& g, const std::vector<int>& part) {
  int64_t cut = 0;
  for (int v = 0; v < g.n; ++v)
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
      if (part[g.adj[k]] != part[v]) cut += g.ew[k];
  return cut / 2;
}

struct HeapEntry {
  int gain;
  int v;
  bool operator<(const HeapEntry& o) const { return gain < o.gain || (gain == o.gain && v > o.v); }
};

// Boundary Fiduccia-Mattheyses refinement of a 2-way partition with a
// balance constraint (side s may hold at most maxw[s]).
void fm_refine(const Graph& g, std::vector<int>& part, const int64_t maxw[2], int npasses) {
  const int n = g.n;
  if (n < 2) return;
  std::vector<int> id(n, 0), ed(n, 0);
  int64_t pw[2] = {0, 0};
  for (int v = 0; v < n; ++v) {
    pw[part[v]] += g.vw[v];
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      if (part[g.adj[k]] == part[v]) id[v] += g.ew[k];
      else ed[v] += g.ew[k];
    }
  }
  int64_t cut = 0;
  for (int v = 0; v < n; ++v) cut += ed[v];
  cut /= 2;
  auto violation = [&](const int64_t w[2]) {
    return std::max<int64_t>(0, w[0] - maxw[0]) + std::max<int64_t>(0, w[1] - maxw[1]);
  };
  std::vector<char> locked(n, 0);
  std::vector<int> moves;
  const int limit = std::max(64, std::min(n / 50, 1000));
  for (int pass = 0; pass < npasses; ++pass) {
    std::priority_queue<HeapEntry> heap[2];
    for (int v = 0; v < n; ++v)
      if (ed[v] > 0 || violation(pw) > 0) heap[part[v]].push({ed[v] - id[v], v});
    moves.clear();
    const int64_t cut0 = cut, viol0 = violation(pw);
    int64_t best_cut = cut, best_viol = viol0;
    size_t best_len = 0;
    int since = 0;
    for (;;) {
      // pick the side to move from
      int from = -1;
      if (pw[0] > maxw[0]) from = 0;
      else if (pw[1] > maxw[1]) from = 1;
      // drop stale heap tops
      for (int s = 0; s < 2; ++s) {
        while (!heap[s].empty()) {
          const HeapEntry& e = heap[s].top();
          if (locked[e.v] || part[e.v] != s || e.gain != ed[e.v] - id[e.v]) { heap[s].pop(); continue; }
          break;
        }
      }
      if (from < 0) {
        int bestg = INT32_MIN;
        for (int s = 0; s < 2; ++s) {
          if (heap[s].empty()) continue;
          const HeapEntry& e = heap[s].top();
          if (pw[1 - s] + g.vw[e.v] > maxw[1 - s]) continue;
          if (e.gain > bestg) { bestg = e.gain; from = s; }
        }
        if (from < 0) break;
      }
      if (heap[from].empty()) break;
      HeapEntry e = heap[from].top();
      heap[from].pop();
      int v = e.v, to = 1 - from;
      // move v
      part[v] = to;
      pw[from] -= g.vw[v];
      pw[to] += g.vw[v];
      cut -= (ed[v] - id[v]);
      std::swap(id[v], ed[v]);
      locked[v] = 1;
      moves.push_back(v);
      for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        int u = g.adj[k], w = g.ew[k];
        if (part[u] == to) { id[u] += w; ed[u] -= w; }
        else { ed[u] += w; id[u] -= w; }
        if (!locked[u]) heap[part[u]].push({ed[u] - id[u], u});
      }
      int64_t viol = violation(pw);
      if (viol < best_viol || (viol == best_viol && cut < best_cut)) {
        best_viol = viol;
        best_cut = cut;
        best_len = moves.size();
        since = 0;
      } else if (++since > limit) {
        break;
      }
    }
    // roll back to the best prefix
    for (size_t m = moves.size(); m > best_len; --m) {
      int v = moves[m - 1], from = part[v], to = 1 - from;
      part[v] = to;
      pw[from] -= g.vw[v];
      pw[to] += g.vw[v];
      cut -= (ed[v] - id[v]);
      std::swap(id[v], ed[v]);
      for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        int u = g.adj[k], w = g.ew[k];
        if (part[u] == to) { id[u] += w; ed[u] -= w; }
        else { ed[u] += w; id[u] -= w; }
      }
    }
    for (int v : moves) locked[v] = 0;
    if (!(best_viol < viol0 || (best_viol == viol0 && best_cut < cut0))) break;
  }
}

// Greedy graph growing from a seed vertex until side 0 reaches its target.
void grow_bisection(const Graph& g, int64_t tw0, uint64_t& rng, std::vector<int>& part) {
  const int n = g.n;
  part.assign(n, 1);
  std::vector<int> conn0(n, 0), deg(n, 0);
  for (int v = 0; v < n; ++v)
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) deg[v] += g.ew[k];
  std::priority_queue<HeapEntry> pq;
  std::vector<char> inq(n, 0);
  int64_t w0 = 0;
  int next_scan = 0;
  std::vector<int> order;
  random_perm(n, rng, order);
  while (w0 < tw0) {
    int v = -1;
    while (!pq.empty()) {
      HeapEntry e = pq.top();
      pq.pop();
      if (part[e.v] == 0 || e.gain != 2 * conn0[e.v] - deg[e.v]) continue;
      v = e.v;
      break;
    }
    if (v < 0) {
      while (next_scan < n && part[order[next_scan]] == 0) ++next_scan;
      if (next_scan >= n) break;
      v = order[next_scan];
    }
    part[v] = 0;
    w0 += g.vw[v];
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int u = g.adj[k];
      if (part[u] == 0) continue;
      conn0[u] += g.ew[k];
      pq.push({2 * conn0[u] - deg[u], u});
    }
  }
}

}  // namespace

Graph graph_from_pattern(int n, const int* ia, const int* ja) {
  Graph g;
  g.n = n;
  std::vector<int> cnt(n + 1, 0);
  for (int i = 0; i < n; ++i)
    for (int k = ia[i]; k < ia[i + 1]; ++k) {
      int j = ja[k];
      if (j == i) continue;
      cnt[i + 1]++;
      cnt[j + 1]++;
    }
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> tmp(cnt[n]);
  std::vector<int> pos(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int k = ia[i]; k < ia[i + 1]; ++k) {
      int j = ja[k];
      if (j == i) continue;
      tmp[pos[i]++] = j;
      tmp[pos[j]++] = i;
    }
  g.xadj.assign(n + 1, 0);
  g.adj.reserve(cnt[n]);
  for (int i = 0; i < n; ++i) {
    auto b = tmp.begin() + cnt[i], e = tmp.begin() + cnt[i + 1];
    std::sort(b, e);
    auto last = std::unique(b, e);
    g.adj.insert(g.adj.end(), b, last);
    g.xadj[i + 1] = (int)g.adj.size();
  }
  g.ew.assign(g.adj.size(), 1);
  g.vw.assign(n, 1);
  return g;
}

Graph graph_induced(const Graph& g, const std::vector<int>& nodes, std::vector<int>& mark) {
  if ((int)mark.size() < g.n) mark.assign(g.n, -1);
  for (int i = 0; i < (int)nodes.size(); ++i) mark[nodes[i]] = i;
  Graph s;
  s.n = (int)nodes.size();
  s.xadj.assign(s.n + 1, 0);
  s.vw.resize(s.n);
  for (int i = 0; i < s.n; ++i) {
    int v = nodes[i];
    s.vw[i] = g.vw[v];
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int u = mark[g.adj[k]];
      if (u >= 0) { s.adj.push_back(u); s.ew.push_back(g.ew[k]); }
    }
    s.xadj[i + 1] = (int)s.adj.size();
  }
  for (int v : nodes) mark[v] = -1;
  return s;
}

std::vector<std::vector<int>> graph_components(const Graph& g) {
  std::vector<std::vector<int>> comps;
  std::vector<char> seen(g.n, 0);
  for (int s = 0; s < g.n; ++s) {
    if (seen[s]) continue;
    comps.emplace_back();
    auto& c = comps.back();
    c.push_back(s);
    seen[s] = 1;
    for (size_t h = 0; h < c.size(); ++h) {
      int v = c[h];
      for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
        int u = g.adj[k];
        if (!seen[u]) { seen[u] = 1; c.push_back(u); }
      }
    }
  }
  return comps;
}

int64_t graph_bisect(const Graph& g, double frac0, uint64_t seed, std::vector<int>& part) {
  const int n = g.n;
  part.assign(n, 0);
  if (n < 2) return 0;
  const double ubf = 0.03;
  uint64_t rng = seed;
  // coarsening hierarchy
  std::vector<Graph> levels;
  std::vector<std::vector<int>> cmaps;
  const Graph* cur = &g;
  while (cur->n > 160) {
    std::vector<int> cmap;
    Graph c = coarsen(*cur, rng, cmap);
    if (c.n > 0.92 * cur->n) break;
    cmaps.push_back(std::move(cmap));
    levels.push_back(std::move(c));
    cur = &levels.back();
  }
  const int64_t W = g.total_vw();
  const int64_t tw0 = (int64_t)(frac0 * (double)W + 0.5), tw1 = W - tw0;
  // initial partition on the coarsest graph: best of several grown bisections
  const Graph& cg = *cur;
  int maxvw = 1;
  for (int v = 0; v < cg.n; ++v) maxvw = std::max(maxvw, cg.vw[v]);
  int64_t maxw[2] = {std::max<int64_t>((int64_t)(tw0 * (1.0 + ubf)), tw0 + maxvw),
                     std::max<int64_t>((int64_t)(tw1 * (1.0 + ubf)), tw1 + maxvw)};
  std::vector<int> best, trial;
  int64_t best_cut = -1, best_viol = -1;
  for (int t = 0; t < 8; ++t) {
    grow_bisection(cg, tw0, rng, trial);
    fm_refine(cg, trial, maxw, 6);
    int64_t pw[2] = {0, 0};
    for (int v = 0; v < cg.n; ++v) pw[trial[v]] += cg.vw[v];
    int64_t viol = std::max<int64_t>(0, pw[0] - maxw[0]) + std::max<int64_t>(0, pw[1] - maxw[1]);
    int64_t c = edge_cut(cg, trial);
    if (best_cut < 0 || viol < best_viol || (viol == best_viol && c < best_cut)) {
      best_cut = c;
      best_viol = viol;
      best = trial;
    }
  }
  // uncoarsen + refine
  std::vector<int> p = best;
  for (int l = (int)levels.size() - 1; l >= 0; --l) {
    const Graph& fine = (l == 0) ? g : levels[l - 1];
    const std::vector<int>& cmap = cmaps[l];
    std::vector<int> fp(fine.n);
    for (int v = 0; v < fine.n; ++v) fp[v] = p[cmap[v]];
    p.swap(fp);
    int mv = 1;
    for (int v = 0; v < fine.n; ++v) mv = std::max(mv, fine.vw[v]);
    int64_t mw[2] = {std::max<int64_t>((int64_t)(tw0 * (1.0 + ubf)), tw0 + mv),
                     std::max<int64_t>((int64_t)(tw1 * (1.0 + ubf)), tw1 + mv)};
    fm_refine(fine, p, mw, 4);
  }
  part.swap(p);
  return edge_cut(g, part);
}

namespace {
void kway_rec(const Graph& g, const std::vector<int>& nodes, int k, int base, uint64_t seed, std::vector<int>& part,
              std::vector<int>& mark) {
  if (k <= 1 || nodes.size() <= 1) {
    for (int v : nodes) part[v] = base;
    return;
  }
  Graph s = graph_induced(g, nodes, mark);
  int k0 = k / 2, k1 = k - k0;
  std::vector<int> p;
  graph_bisect(s, (double)k0 / (double)k, seed, p);
  std::vector<int> n0, n1;
  for (int i = 0; i < s.n; ++i) (p[i] == 0 ? n0 : n1).push_back(nodes[i]);
  kway_rec(g, n0, k0, base, derive_seed(seed, 2 * base + 1), part, mark);
  kway_rec(g, n1, k1, base + k0, derive_seed(seed, 2 * base + 2), part, mark);
}
}  // namespace

void graph_kway(const Graph& g, int k, uint64_t seed, std::vector<int>& part) {
  part.assign(g.n, 0);
  std::vector<int> all(g.n), mark;
  for (int i = 0; i < g.n; ++i) all[i] = i;
  kway_rec(g, all, k, 0, seed, part, mark);
}

int graph_vertex_separator(const Graph& g, const std::vector<int>& part, std::vector<char>& sep) {
  const int n = g.n;
  sep.assign(n, 0);
  std::vector<int> idx(n, -1), A, B;
  for (int v = 0; v < n; ++v) {
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
      if (part[g.adj[k]] != part[v]) {
        if (part[v] == 0) { idx[v] = (int)A.size(); A.push_back(v); }
        else { idx[v] = (int)B.size(); B.push_back(v); }
        break;
      }
  }
  const int na = (int)A.size(), nb = (int)B.size();
  if (na == 0 || nb == 0) return 0;
  // bipartite adjacency A -> B
  std::vector<int> bx(na + 1, 0), badj;
  for (int a = 0; a < na; ++a) {
    int v = A[a];
    for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k) {
      int u = g.adj[k];
      if (part[u] == 1 && idx[u] >= 0) badj.push_back(idx[u]);
    }
    bx[a + 1] = (int)badj.size();
  }
  std::vector<int> mateA(na, -1), mateB(nb, -1);
  for (int a = 0; a < na; ++a)  // greedy start
    for (int k = bx[a]; k < bx[a + 1]; ++k)
      if (mateB[badj[k]] < 0) { mateA[a] = badj[k]; mateB[badj[k]] = a; break; }
  // augmenting paths by BFS over alternating trees
  std::vector<int> stampB(nb, -1), parentB(nb, -1), queue;
  for (int root = 0; root < na; ++root) {
    if (mateA[root] >= 0) continue;
    queue.clear();
    queue.push_back(root);
    int found = -1;
    for (size_t h = 0; h < queue.size() && found < 0; ++h) {
      int a = queue[h];
      for (int k = bx[a]; k < bx[a + 1]; ++k) {
        int b = badj[k];
        if (stampB[b] == root) continue;
        stampB[b] = root;
        parentB[b] = a;
        if (mateB[b] < 0) { found = b; break; }
        queue.push_back(mateB[b]);
      }
    }
    for (int b = found; b >= 0;) {
      int a = parentB[b];
      int nb2 = mateA[a];
      mateA[a] = b;
      mateB[b] = a;
      b = nb2;
    }
  }
  // Koenig: Z = reachable from unmatched A by alternating paths
  std::vector<char> zA(na, 0), zB(nb, 0);
  queue.clear();
  for (int a = 0; a < na; ++a)
    if (mateA[a] < 0) { zA[a] = 1; queue.push_back(a); }
  for (size_t h = 0; h < queue.size(); ++h) {
    int a = queue[h];
    for (int k = bx[a]; k < bx[a + 1]; ++k) {
      int b = badj[k];
      if (zB[b]) continue;
      zB[b] = 1;
      int a2 = mateB[b];
      if (a2 >= 0 && !zA[a2]) { zA[a2] = 1; queue.push_back(a2); }
    }
  }
  int cnt = 0;
  for (int a = 0; a < na; ++a)
    if (!zA[a]) { sep[A[a]] = 1; ++cnt; }
  for (int b = 0; b < nb; ++b)
    if (zB[b]) { sep[B[b]] = 1; ++cnt; }
  return cnt;
}

namespace {

using LevelSets = std::vector<std::vector<std::vector<int>>>;

// Nested dissection level structure (SetupPermutationNDRecursive semantics):
// ls[0] holds the leaves, ls[j] the separators found j levels above them.
void nd_rec(const Graph& g, int clvl, int& tlvl, int minsep, uint64_t seed, LevelSets& ls) {
  ls.clear();
  if (tlvl < 2 || g.n < 2) {
    tlvl = 1;
    ls.assign(1, std::vector<std::vector<int>>(1));
    ls[0][0].resize(g.n);
    for (int i = 0; i < g.n; ++i) ls[0][0][i] = i;
    return;
  }
  auto comps = graph_components(g);
  if (clvl >= tlvl - 1) {
    tlvl = 1;
    ls.assign(1, std::move(comps));
    return;
  }
  ls.resize(tlvl - clvl);
  int tl_max = 1;
  std::vector<int> mark;
  for (size_t ci = 0; ci < comps.size(); ++ci) {
    std::vector<int>& comp = comps[ci];
    const int nS = (int)comp.size();
    if (nS <= minsep || nS <= 2) {
      ls[0].push_back(comp);
      continue;
    }
    Graph C = graph_induced(g, comp, mark);
    std::vector<int> part;
    graph_bisect(C, 0.5, derive_seed(seed, 1000003ULL * (uint64_t)clvl + ci), part);
    std::vector<char> sep;
    int nsep = graph_vertex_separator(C, part, sep);
    int interior[2] = {0, 0};
    for (int v = 0; v < C.n; ++v)
      if (!sep[v]) interior[part[v]]++;
    if (nsep == 0 || interior[0] == 0 || interior[1] == 0) {
      ls[0].push_back(comp);
      continue;
    }
    std::vector<int> keep, sepn;
    for (int v = 0; v < C.n; ++v) (sep[v] ? sepn : keep).push_back(v);
    Graph Bg = graph_induced(C, keep, mark);
    int tl = tlvl;
    LevelSets sub;
    nd_rec(Bg, clvl + 1, tl, minsep, derive_seed(seed, 7919ULL * (uint64_t)(clvl + 1) + ci), sub);
    for (int j = 0; j < tl; ++j)
      for (auto& dom : sub[j]) {
        std::vector<int> mapped(dom.size());
        for (size_t t = 0; t < dom.size(); ++t) mapped[t] = comp[keep[dom[t]]];
        ls[j].push_back(std::move(mapped));
      }
    std::vector<int> sepg(sepn.size());
    for (size_t t = 0; t < sepn.size(); ++t) sepg[t] = comp[sepn[t]];
    if ((int)ls.size() <= tl) ls.resize(tl + 1);
    ls[tl].push_back(std::move(sepg));
    tl_max = std::max(tl_max, tl + 1);
  }
  tlvl = tl_max;
  ls.resize(tlvl);
}

void single_block(int clvl, int& tlvl, std::vector<int>& map, std::vector<int>& mapptr) {
  int domi = mapptr[clvl];
  mapptr[clvl + 1] = domi + 1;
  std::fill(map.begin(), map.end(), domi);
  tlvl = clvl + 1;
}

void rkway_rec(const Graph& g, bool vertexsep, int clvl, int& tlvl, int num_dom, const RkwayParams& prm,
               std::vector<int>& map, std::vector<int>& mapptr, uint64_t seed) {
  const int nA = g.n;
  map.assign(nA, 0);
  if (tlvl < 2 || nA < 2) {
    mapptr[clvl + 1] = mapptr[clvl] + 1;
    tlvl = clvl + 1;
    std::fill(map.begin(), map.end(), mapptr[clvl]);
    return;
  }
  if (clvl == tlvl - 1) { single_block(clvl, tlvl, map, mapptr); return; }
  if (prm.minsep >= nA || num_dom >= nA) { single_block(clvl, tlvl, map, mapptr); return; }

  std::vector<int> lmap(nA, 0);
  std::vector<char> sep(nA, 0);
  int num_dom2 = 0, edgecut = 0;
  if (vertexsep) {
    if ((num_dom & (num_dom - 1)) != 0) throw std::invalid_argument("vertex-separator partition needs a power-of-two subdomain count");
    int nd_tlvl = 0;
    for (int t = num_dom; t > 0; t >>= 1) ++nd_tlvl;
    LevelSets ls;
    nd_rec(g, 0, nd_tlvl, 2, seed, ls);
    num_dom2 = (int)ls[0].size();
    std::fill(lmap.begin(), lmap.end(), num_dom2 - 1);
    std::fill(sep.begin(), sep.end(), 1);
    int ninterior = 0;
    for (int d = 0; d < num_dom2; ++d)
      for (int v : ls[0][d]) { lmap[v] = d; sep[v] = 0; ++ninterior; }
    edgecut = nA - ninterior;
  } else {
    std::vector<int> part;
    graph_kway(g, num_dom, seed, part);
    std::vector<int> cnt(num_dom, 0), remap(num_dom, -1);
    for (int v = 0; v < nA; ++v) cnt[part[v]]++;
    for (int d = 0; d < num_dom; ++d)
      if (cnt[d] > 0) remap[d] = num_dom2++;
    for (int v = 0; v < nA; ++v) lmap[v] = remap[part[v]];
    std::vector<int> sep_size(num_dom2, 0), dsize(num_dom2, 0);
    for (int v = 0; v < nA; ++v) {
      dsize[lmap[v]]++;
      for (int k = g.xadj[v]; k < g.xadj[v + 1]; ++k)
        if (lmap[g.adj[k]] != lmap[v]) { sep[v] = 1; sep_size[lmap[v]]++; ++edgecut; break; }
    }
    for (int d = 0; d < num_dom2; ++d)
      if (sep_size[d] == dsize[d]) { edgecut = nA; break; }
  }
  const int domi = mapptr[clvl];
  if (edgecut == 0) {
    mapptr[clvl + 1] = domi + num_dom2;
    for (int v = 0; v < nA; ++v) map[v] = lmap[v] + domi;
    tlvl = clvl + 1;
    return;
  }
  if (num_dom2 < num_dom || edgecut == nA) { single_block(clvl, tlvl, map, mapptr); return; }
  std::vector<int> cnodes;
  cnodes.reserve(edgecut);
  for (int v = 0; v < nA; ++v)
    if (sep[v]) cnodes.push_back(v);
  std::vector<int> mark;
  Graph C = graph_induced(g, cnodes, mark);
  mapptr[clvl + 1] = domi + num_dom2;
  int nd_next = num_dom;
  if (nd_next > prm.kmin) {
    nd_next = nd_next / std::max(1, prm.kfactor);
    if (nd_next < prm.kmin) nd_next = prm.kmin;
  }
  std::vector<int> mapc;
  rkway_rec(C, false, clvl + 1, tlvl, nd_next, prm, mapc, mapptr, derive_seed(seed, 31 + clvl));
  int j = 0;
  for (int v = 0; v < nA; ++v) map[v] = sep[v] ? mapc[j++] : lmap[v] + domi;
}

}  // namespace

void rkway_levels(const Graph& g, const RkwayParams& prm, int& nlev, std::vector<int>& map, std::vector<int>& mapptr) {
  int tlvl = prm.nlev;
  mapptr.assign(std::max(tlvl, 1) + 1, 0);
  int num_dom = std::min(prm.ncomp, g.n);
  if (num_dom < 2 && tlvl >= 2 && g.n >= 2) num_dom = 2;
  rkway_rec(g, prm.vertexsep, 0, tlvl, num_dom, prm, map, mapptr, prm.seed);
  nlev = tlvl;
  mapptr.resize(tlvl + 1);
}

}  // namespace pgs
