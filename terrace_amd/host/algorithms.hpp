// The algorithms of the reference's benchmark, restated over the graphstore
// shim (include/terrace/graphstore.hpp).  Same control flow, functor names,
// functor constructor arguments and VertexSubset constructors as
// include/BFS.h, Pagerank.h, Components.h, BC.h, TC.h and
// BellmanFordUnweighted.h; the per-edge work runs on the GPU through the
// functors' device ops, so the functors hold device pointers (DeviceArray
// converts to them) and the element-wise initialisation loops become
// DeviceArray fills.  Results are malloc'ed host arrays the caller frees, as
// in the reference.
#pragma once

#include <cstdlib>
#include <numeric>
#include <vector>

#include "../../include/terrace/graphstore.hpp"

namespace graphstore {

// ---- BFS (include/BFS.h:27-84) ------------------------------------------
struct BFS_F {
  int32_t *Parents;
  BFS_F(int32_t *_Parents) : Parents(_Parents) {}
};
template <>
struct edge_op_traits<BFS_F> {
  static constexpr int op = TG_OP_BFS;
  static tg_op_args args(const BFS_F &f) {
    tg_op_args a{};
    a.parent = f.Parents;
    return a;
  }
};

inline int32_t *BFS_with_edge_map(Graph &G, uint32_t start) {
  const uint64_t n = G.get_num_vertices();
  DeviceArray<int32_t> Parents(n);
  Parents.fill(-1);
  Parents.set(start, static_cast<int32_t>(start));
  VertexSubset frontier = VertexSubset(start, n);  // creates initial frontier
  while (frontier.not_empty()) {                   // loop until frontier is empty
    VertexSubset next_frontier = edgeMap(G, frontier, BFS_F(Parents), true);
    frontier.del();
    frontier = next_frontier;
  }
  frontier.del();
  int32_t *out = static_cast<int32_t *>(malloc(n * sizeof(int32_t)));
  Parents.download(out, n);
  return out;
}

// ---- PageRank (include/Pagerank.h:32-184), undamped ---------------------
struct PR_F {
  double *p_curr, *p_next;
  Graph &G;
  PR_F(double *_p_curr, double *_p_next, Graph &_G) : p_curr(_p_curr), p_next(_p_next), G(_G) {}
};
template <>
struct edge_op_traits<PR_F> {
  static constexpr int op = TG_OP_PR;
  static tg_op_args args(const PR_F &f) {
    tg_op_args a{};
    a.src_val = f.p_curr;
    a.val = f.p_next;
    return a;
  }
};
struct PR_Vertex {
  double *p_curr;
  Graph &G;
  PR_Vertex(double *_p_curr, Graph &_G) : p_curr(_p_curr), G(_G) {}
};
template <>
struct vertex_op_traits<PR_Vertex> {
  static constexpr int op = TG_VOP_PR_VERTEX;
  static tg_vmap_args args(const PR_Vertex &f) {
    tg_vmap_args a{};
    a.val = f.p_curr;
    return a;
  }
};
struct PR_Vertex_Reset {
  double *p;
  PR_Vertex_Reset(double *_p) : p(_p) {}
};
template <>
struct vertex_op_traits<PR_Vertex_Reset> {
  static constexpr int op = TG_VOP_PR_RESET;
  static tg_vmap_args args(const PR_Vertex_Reset &f) {
    tg_vmap_args a{};
    a.val = f.p;
    return a;
  }
};

inline double *PR_S(Graph &G, long maxIters) {
  const uint64_t n = G.get_num_vertices();
  const double one_over_n = 1 / static_cast<double>(n);
  DeviceArray<double> p_curr(n), p_next(n);
  p_curr.fill(one_over_n);
  VertexSubset Frontier = VertexSubset(0, n, true);
  long iter = 0;
  while (iter++ < maxIters) {
    vertexMap(Frontier, PR_Vertex(p_curr, G), false);
    vertexMap(Frontier, PR_Vertex_Reset(p_next), false);
    edgeMap(G, Frontier, PR_F(p_curr, p_next, G), false);
    std::swap(p_curr, p_next);
  }
  Frontier.del();
  double *out = static_cast<double *>(malloc(n * sizeof(double)));
  p_curr.download(out, n);
  return out;
}

// ---- Connected components (include/Components.h:29-120) -----------------
struct CC_F {
  uint32_t *IDs, *prevIDs;
  CC_F(uint32_t *_IDs, uint32_t *_prevIDs) : IDs(_IDs), prevIDs(_prevIDs) {}
};
template <>
struct edge_op_traits<CC_F> {
  static constexpr int op = TG_OP_CC;
  static tg_op_args args(const CC_F &f) {
    tg_op_args a{};
    a.ids = f.IDs;
    a.prev = f.prevIDs;
    return a;
  }
};
struct CC_Vertex_F {
  uint32_t *IDs, *prevIDs;
  CC_Vertex_F(uint32_t *_IDs, uint32_t *_prevIDs) : IDs(_IDs), prevIDs(_prevIDs) {}
};
template <>
struct vertex_op_traits<CC_Vertex_F> {
  static constexpr int op = TG_VOP_CC_VERTEX;
  static tg_vmap_args args(const CC_Vertex_F &f) {
    tg_vmap_args a{};
    a.ids = f.IDs;
    a.prev = f.prevIDs;
    return a;
  }
};

inline uint32_t *CC(Graph &G) {
  const uint64_t n = G.get_num_vertices();
  DeviceArray<uint32_t> IDs(n), prevIDs(n);
  {
    std::vector<uint32_t> h(n);
    std::iota(h.begin(), h.end(), 0u);
    IDs.upload(h.data(), n);
  }
  VertexSubset Active = VertexSubset(0, n, true);  // initial frontier contains all vertices
  while (Active.not_empty()) {                      // iterate until IDS converge
    vertexMap(Active, CC_Vertex_F(IDs, prevIDs), false);
    VertexSubset next = edgeMap(G, Active, CC_F(IDs, prevIDs), true);
    Active.del();
    Active = next;
  }
  Active.del();
  uint32_t *out = static_cast<uint32_t *>(malloc(n * sizeof(uint32_t)));
  IDs.download(out, n);
  return out;
}

// ---- Betweenness centrality from one source (include/BC.h:34-158) -------
struct BC_F {
  double *NumPaths;
  bool *Visited;
  BC_F(double *_NumPaths, bool *_Visited) : NumPaths(_NumPaths), Visited(_Visited) {}
};
template <>
struct edge_op_traits<BC_F> {
  static constexpr int op = TG_OP_BC_FWD;
  static tg_op_args args(const BC_F &f) {
    tg_op_args a{};
    a.val = f.NumPaths;
    a.visited = reinterpret_cast<const uint8_t *>(f.Visited);
    return a;
  }
};
struct BC_Back_F {
  double *Dependencies;
  bool *Visited;
  BC_Back_F(double *_Dependencies, bool *_Visited) : Dependencies(_Dependencies), Visited(_Visited) {}
};
template <>
struct edge_op_traits<BC_Back_F> {
  static constexpr int op = TG_OP_BC_BACK;
  static tg_op_args args(const BC_Back_F &f) {
    tg_op_args a{};
    a.val = f.Dependencies;
    a.visited = reinterpret_cast<const uint8_t *>(f.Visited);
    return a;
  }
};
struct BC_Vertex_F {
  bool *Visited;
  BC_Vertex_F(bool *_Visited) : Visited(_Visited) {}
};
template <>
struct vertex_op_traits<BC_Vertex_F> {
  static constexpr int op = TG_VOP_BC_VERTEX;
  static tg_vmap_args args(const BC_Vertex_F &f) {
    tg_vmap_args a{};
    a.visited = reinterpret_cast<uint8_t *>(f.Visited);
    return a;
  }
};
struct BC_Back_Vertex_F {
  bool *Visited;
  double *Dependencies, *inverseNumPaths;
  BC_Back_Vertex_F(bool *_Visited, double *_Dependencies, double *_inverseNumPaths)
      : Visited(_Visited), Dependencies(_Dependencies), inverseNumPaths(_inverseNumPaths) {}
};
template <>
struct vertex_op_traits<BC_Back_Vertex_F> {
  static constexpr int op = TG_VOP_BC_BACK_VERTEX;
  static tg_vmap_args args(const BC_Back_Vertex_F &f) {
    tg_vmap_args a{};
    a.visited = reinterpret_cast<uint8_t *>(f.Visited);
    a.val = f.Dependencies;
    a.inv = f.inverseNumPaths;
    return a;
  }
};
// The element-wise parallel_for loops of BC.h (:126-128, :130-132,
// :140-142) as vertex ops over all vertices.
struct Invert_F {
  double *v;
  Invert_F(double *_v) : v(_v) {}
};
template <>
struct vertex_op_traits<Invert_F> {
  static constexpr int op = TG_VOP_INVERT;
  static tg_vmap_args args(const Invert_F &f) {
    tg_vmap_args a{};
    a.val = f.v;
    return a;
  }
};
struct ClearVisited_F {
  bool *Visited;
  ClearVisited_F(bool *_Visited) : Visited(_Visited) {}
};
template <>
struct vertex_op_traits<ClearVisited_F> {
  static constexpr int op = TG_VOP_CLEAR_VISITED;
  static tg_vmap_args args(const ClearVisited_F &f) {
    tg_vmap_args a{};
    a.visited = reinterpret_cast<uint8_t *>(f.Visited);
    return a;
  }
};
struct BC_Final_F {
  double *Dependencies, *inverseNumPaths;
  BC_Final_F(double *_Dependencies, double *_inverseNumPaths)
      : Dependencies(_Dependencies), inverseNumPaths(_inverseNumPaths) {}
};
template <>
struct vertex_op_traits<BC_Final_F> {
  static constexpr int op = TG_VOP_BC_FINAL;
  static tg_vmap_args args(const BC_Final_F &f) {
    tg_vmap_args a{};
    a.val = f.Dependencies;
    a.inv = f.inverseNumPaths;
    return a;
  }
};

inline double *BC(Graph &G, uint32_t start) {
  const uint64_t n = G.get_num_vertices();
  DeviceArray<double> NumPaths(n), Dependencies(n);
  DeviceArray<bool> Visited(n);
  NumPaths.fill(0.0);
  Visited.fill(false);
  Visited.set(start, true);
  NumPaths.set(start, 1.0);
  VertexSubset Frontier = VertexSubset(start, n);  // creates initial frontier
  std::vector<VertexSubset> Levels;
  Levels.push_back(Frontier);
  long round = 0;
  while (Frontier.not_empty()) {  // first phase
    round++;
    VertexSubset output = edgeMap(G, Frontier, BC_F(NumPaths, Visited));
    Frontier = output;  // shallow copy, as in the reference
    Levels.push_back(Frontier);
    vertexMap(Frontier, BC_Vertex_F(Visited), false);  // mark visited
  }
  Dependencies.fill(0.0);
  VertexSubset All = VertexSubset(0, n, true);
  vertexMap(All, Invert_F(NumPaths), false);  // invert numpaths
  Levels[round].del();
  vertexMap(All, ClearVisited_F(Visited), false);  // reuse Visited
  vertexMap(Levels[round - 1], BC_Back_Vertex_F(Visited, Dependencies, NumPaths), false);
  for (long r = round - 2; r >= 0; r--) {  // backwards phase
    edgeMap(G, Levels[r + 1], BC_Back_F(Dependencies, Visited), false);
    Levels[r + 1].del();
    vertexMap(Levels[r], BC_Back_Vertex_F(Visited, Dependencies, NumPaths), false);
  }
  vertexMap(All, BC_Final_F(Dependencies, NumPaths), false);  // the last loop of BC.h
  Levels[0].del();
  All.del();
  double *out = static_cast<double *>(malloc(n * sizeof(double)));
  Dependencies.download(out, n);
  return out;
}

// ---- SSSP_BF with unit lengths (include/BellmanFordUnweighted.h:29-130) --
// Visited is one byte per vertex (the reference uses int), claimed in
// 4-byte words, so the array is padded to a multiple of 4 bytes.
struct BF_F {
  int32_t *ShortestPathLen;
  uint8_t *Visited;
  BF_F(int32_t *_ShortestPathLen, uint8_t *_Visited) : ShortestPathLen(_ShortestPathLen), Visited(_Visited) {}
};
template <>
struct edge_op_traits<BF_F> {
  static constexpr int op = TG_OP_BF;
  static tg_op_args args(const BF_F &f) {
    tg_op_args a{};
    a.dist = f.ShortestPathLen;
    a.visited = f.Visited;
    return a;
  }
};
struct BF_Vertex_F {
  uint8_t *Visited;
  BF_Vertex_F(uint8_t *_Visited) : Visited(_Visited) {}
};
template <>
struct vertex_op_traits<BF_Vertex_F> {
  static constexpr int op = TG_VOP_CLEAR_VISITED;
  static tg_vmap_args args(const BF_Vertex_F &f) {
    tg_vmap_args a{};
    a.visited = f.Visited;
    return a;
  }
};

// On a weighted graph the same loop is BellmanFord.h's SSSP_BF (:67-143):
// the device op takes the stored weights as edge lengths.
inline int32_t *SSSP_BF(Graph &G, long start) {
  const uint64_t n = G.get_num_vertices();
  const int32_t inf = 0x7FFFFFFF / 2;
  DeviceArray<int32_t> ShortestPathLen(n);
  ShortestPathLen.fill(inf);
  ShortestPathLen.set(static_cast<uint64_t>(start), 0);
  DeviceArray<uint8_t> Visited((n + 4) & ~3ull);
  Visited.fill(0);
  VertexSubset Frontier = VertexSubset(static_cast<uint32_t>(start), n);  // initial frontier
  uint64_t round = 0;
  while (Frontier.not_empty()) {
    if (round == n) {  // negative weight cycle rule of the reference
      ShortestPathLen.fill(-inf);
      break;
    }
    // sparse only in the reference (threshold 1)
    VertexSubset output = edgeMap(G, Frontier, BF_F(ShortestPathLen, Visited), true, 1);
    vertexMap(output, BF_Vertex_F(Visited), false);
    Frontier.del();
    Frontier = output;
    round++;
  }
  Frontier.del();
  int32_t *out = static_cast<int32_t *>(malloc(n * sizeof(int32_t)));
  ShortestPathLen.download(out, n);
  return out;
}

// ---- Triangle counting (include/TC.h:51-87) ------------------------------
// countF's update (TC.h:51-70) counts count_common(s, d) for s > d; the
// device op sums it over the frontier's edges into counts[0] (one device
// counter instead of one per worker).
struct countF {
  Graph &G;
  uint64_t *counts;
  countF(Graph &G_, uint64_t *_counts) : G(G_), counts(_counts) {}
};
template <>
struct edge_op_traits<countF> {
  static constexpr int op = TG_OP_TC;
  static tg_op_args args(const countF &f) {
    tg_op_args a{};
    a.count = f.counts;
    return a;
  }
};

inline uint64_t TC(Graph &G) {
  const uint32_t n = G.get_num_vertices();
  DeviceArray<uint64_t> counts(1);
  counts.fill(0);
  VertexSubset Frontier(0, n, true);  // frontier contains all vertices
  edgeMap(G, Frontier, countF(G, counts), false);
  Frontier.del();
  return counts.to_host()[0];
}

}  // namespace graphstore
