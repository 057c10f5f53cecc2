// Graph partitioning for the GeMSLR level structure (host setup).
//
// The reference calls METIS_PartGraphKway / METIS_PartGraphRecursive through
// CsrMatrixMetisKwayHost (ParGeMSLR/SRC/matrices/matrixops.cpp:7074-7435) and
// builds levels with SetupPermutationRKwayRecursive (matrixops.cpp:11454-11692)
// and SetupPermutationNDRecursive (matrixops.cpp:11736-11911).  METIS is not
// part of this library: bisection here is a multilevel scheme of its own
// (heavy-edge matching coarsening, greedy graph growing, boundary FM
// refinement), vertex separators come from a maximum bipartite matching and
// Koenig's theorem, and k-way partitions come from recursive bisection.
#pragma once

#include <cstdint>
#include <vector>

namespace pgs {

struct Graph {
  int n = 0;
  std::vector<int> xadj{0}, adj, ew, vw;
  int64_t total_vw() const {
    int64_t s = 0;
    for (int v : vw) s += v;
    return s;
  }
};

// Adjacency of A + A^T without self loops, unit vertex and edge weights.
Graph graph_from_pattern(int n, const int* ia, const int* ja);

// Induced subgraph on `nodes` (local numbering = position in nodes).
Graph graph_induced(const Graph& g, const std::vector<int>& nodes, std::vector<int>& mark_buf);

// Connected components; each component lists local vertex ids in BFS order
// from its smallest vertex.
std::vector<std::vector<int>> graph_components(const Graph& g);

// Two-way partition: part[v] in {0,1}; side 0 targets `frac0` of the weight.
// Returns the weighted edge cut.
int64_t graph_bisect(const Graph& g, double frac0, uint64_t seed, std::vector<int>& part);

// k-way partition by recursive bisection: part[v] in [0, k).
void graph_kway(const Graph& g, int k, uint64_t seed, std::vector<int>& part);

// Minimum vertex separator drawn from the edge-cut boundary of a bisection
// (maximum matching on the boundary bipartite graph + Koenig's theorem).
// sep[v] = 1 marks separator vertices.  Returns the separator size.
int graph_vertex_separator(const Graph& g, const std::vector<int>& part, std::vector<char>& sep);

// Level structure of GeMSLR's recursive k-way partition
// (SetupPermutationRKwayRecursive).  On return map[v] = domain id and
// mapptr[l]..mapptr[l+1] are the domain ids of level l; nlev is updated to the
// number of levels actually built.
struct RkwayParams {
  int nlev = 3;
  int ncomp = 4;
  int minsep = 10;
  int kmin = 1;
  int kfactor = 1;
  bool vertexsep = true;
  uint64_t seed = 0x5eed;
};
void rkway_levels(const Graph& g, const RkwayParams& prm, int& nlev, std::vector<int>& map, std::vector<int>& mapptr);

}  // namespace pgs
