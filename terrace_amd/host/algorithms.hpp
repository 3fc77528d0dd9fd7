// The algorithms of the reference's benchmark, restated over the graphstore
// shim (include/terrace/graphstore.hpp).  Same control flow and functor
// roles as include/BFS.h, Pagerank.h, Components.h, BC.h, TC.h and
// BellmanFordUnweighted.h;
// the per-edge work runs on the GPU through the functors' device ops, and
// the state arrays live in HBM.  Results are malloc'ed host arrays the
// caller frees, as in the reference.
#pragma once

#include <cstdlib>
#include <numeric>
#include <vector>

#include "../../include/terrace/graphstore.hpp"

namespace graphstore {

// ---- BFS (include/BFS.h:27-84) ------------------------------------------
struct BFS_F {
  DeviceArray<int32_t> *Parents;
  uint32_t depth;  // level given to vertices reached in this round
};
template <>
struct edge_op_traits<BFS_F> {
  static constexpr int op = TG_OP_BFS;
  static tg_op_args args(const BFS_F &f) {
    tg_op_args a{};
    a.parent = f.Parents->data();
    a.depth = f.depth;
    return a;
  }
};

inline int32_t *BFS_with_edge_map(Graph &G, uint32_t src) {
  const uint64_t n = G.get_num_vertices();
  DeviceArray<int32_t> Parents(n);
  Parents.fill(-1);
  Parents.set(src, static_cast<int32_t>(src));
  VertexSubset frontier(G, src, n);
  uint32_t depth = 0;
  while (frontier.not_empty()) {
    VertexSubset next = edgeMap(G, frontier, BFS_F{&Parents, ++depth}, true);
    frontier.del();
    frontier = next;
  }
  frontier.del();
  int32_t *out = static_cast<int32_t *>(malloc(n * sizeof(int32_t)));
  Parents.download(out, n);
  return out;
}

// ---- PageRank (include/Pagerank.h:32-184), undamped ---------------------
struct PR_F {
  DeviceArray<double> *p_curr, *p_next;
};
template <>
struct edge_op_traits<PR_F> {
  static constexpr int op = TG_OP_PR;
  static tg_op_args args(const PR_F &f) {
    tg_op_args a{};
    a.src_val = f.p_curr->data();
    a.val = f.p_next->data();
    return a;
  }
};
struct PR_Vertex {
  DeviceArray<double> *p_curr;
};
template <>
struct vertex_op_traits<PR_Vertex> {
  static constexpr int op = TG_VOP_PR_VERTEX;
  static tg_vmap_args args(const PR_Vertex &f) {
    tg_vmap_args a{};
    a.val = f.p_curr->data();
    return a;
  }
};
struct PR_Vertex_Reset {
  DeviceArray<double> *p;
};
template <>
struct vertex_op_traits<PR_Vertex_Reset> {
  static constexpr int op = TG_VOP_PR_RESET;
  static tg_vmap_args args(const PR_Vertex_Reset &f) {
    tg_vmap_args a{};
    a.val = f.p->data();
    return a;
  }
};

inline double *PR_S(Graph &G, long maxIters) {
  const uint64_t n = G.get_num_vertices();
  DeviceArray<double> p_curr(n), p_next(n);
  p_curr.fill(1 / static_cast<double>(n));
  VertexSubset Frontier(G, 0, n, true);
  long iter = 0;
  while (iter++ < maxIters) {
    vertexMap(Frontier, PR_Vertex{&p_curr}, false);
    vertexMap(Frontier, PR_Vertex_Reset{&p_next}, false);
    edgeMap(G, Frontier, PR_F{&p_curr, &p_next}, false);
    std::swap(p_curr, p_next);
  }
  Frontier.del();
  double *out = static_cast<double *>(malloc(n * sizeof(double)));
  p_curr.download(out, n);
  return out;
}

// ---- Connected components (include/Components.h:29-120) -----------------
struct CC_F {
  DeviceArray<uint32_t> *IDs, *prevIDs;
};
template <>
struct edge_op_traits<CC_F> {
  static constexpr int op = TG_OP_CC;
  static tg_op_args args(const CC_F &f) {
    tg_op_args a{};
    a.ids = f.IDs->data();
    a.prev = f.prevIDs->data();
    return a;
  }
};
struct CC_Vertex_F {
  DeviceArray<uint32_t> *IDs, *prevIDs;
};
template <>
struct vertex_op_traits<CC_Vertex_F> {
  static constexpr int op = TG_VOP_CC_VERTEX;
  static tg_vmap_args args(const CC_Vertex_F &f) {
    tg_vmap_args a{};
    a.ids = f.IDs->data();
    a.prev = f.prevIDs->data();
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
  VertexSubset Active(G, 0, n, true);
  while (Active.not_empty()) {
    vertexMap(Active, CC_Vertex_F{&IDs, &prevIDs}, false);
    VertexSubset next = edgeMap(G, Active, CC_F{&IDs, &prevIDs}, true);
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
  DeviceArray<double> *NumPaths;
  DeviceArray<uint8_t> *Visited;
};
template <>
struct edge_op_traits<BC_F> {
  static constexpr int op = TG_OP_BC_FWD;
  static tg_op_args args(const BC_F &f) {
    tg_op_args a{};
    a.val = f.NumPaths->data();
    a.visited = f.Visited->data();
    return a;
  }
};
struct BC_Back_F {
  DeviceArray<double> *Dependencies;
  DeviceArray<uint8_t> *Visited;
};
template <>
struct edge_op_traits<BC_Back_F> {
  static constexpr int op = TG_OP_BC_BACK;
  static tg_op_args args(const BC_Back_F &f) {
    tg_op_args a{};
    a.val = f.Dependencies->data();
    a.visited = f.Visited->data();
    return a;
  }
};
struct BC_Vertex_F {
  DeviceArray<uint8_t> *Visited;
};
template <>
struct vertex_op_traits<BC_Vertex_F> {
  static constexpr int op = TG_VOP_BC_VERTEX;
  static tg_vmap_args args(const BC_Vertex_F &f) {
    tg_vmap_args a{};
    a.visited = f.Visited->data();
    return a;
  }
};
struct BC_Back_Vertex_F {
  DeviceArray<uint8_t> *Visited;
  DeviceArray<double> *Dependencies, *inverseNumPaths;
};
template <>
struct vertex_op_traits<BC_Back_Vertex_F> {
  static constexpr int op = TG_VOP_BC_BACK_VERTEX;
  static tg_vmap_args args(const BC_Back_Vertex_F &f) {
    tg_vmap_args a{};
    a.visited = f.Visited->data();
    a.val = f.Dependencies->data();
    a.inv = f.inverseNumPaths->data();
    return a;
  }
};
// the element-wise loops of BC.h as vertex ops over all vertices
struct Invert_F {
  DeviceArray<double> *v;
};
template <>
struct vertex_op_traits<Invert_F> {
  static constexpr int op = TG_VOP_INVERT;
  static tg_vmap_args args(const Invert_F &f) {
    tg_vmap_args a{};
    a.val = f.v->data();
    return a;
  }
};
struct ClearVisited_F {
  DeviceArray<uint8_t> *Visited;
};
template <>
struct vertex_op_traits<ClearVisited_F> {
  static constexpr int op = TG_VOP_CLEAR_VISITED;
  static tg_vmap_args args(const ClearVisited_F &f) {
    tg_vmap_args a{};
    a.visited = f.Visited->data();
    return a;
  }
};
struct BC_Final_F {
  DeviceArray<double> *Dependencies, *inverseNumPaths;
};
template <>
struct vertex_op_traits<BC_Final_F> {
  static constexpr int op = TG_VOP_BC_FINAL;
  static tg_vmap_args args(const BC_Final_F &f) {
    tg_vmap_args a{};
    a.val = f.Dependencies->data();
    a.inv = f.inverseNumPaths->data();
    return a;
  }
};

inline double *BC(Graph &G, uint32_t start) {
  const uint64_t n = G.get_num_vertices();
  DeviceArray<double> NumPaths(n), Dependencies(n);
  DeviceArray<uint8_t> Visited(n);
  NumPaths.fill(0.0);
  Visited.fill(0);
  Visited.set(start, 1);
  NumPaths.set(start, 1.0);
  VertexSubset Frontier(G, start, n);
  std::vector<VertexSubset> Levels;
  Levels.push_back(Frontier);
  long round = 0;
  while (Frontier.not_empty()) {
    round++;
    VertexSubset output = edgeMap(G, Frontier, BC_F{&NumPaths, &Visited});
    Frontier = output;
    Levels.push_back(Frontier);
    vertexMap(Frontier, BC_Vertex_F{&Visited}, false);
  }
  Dependencies.fill(0.0);
  VertexSubset All(G, 0, n, true);
  vertexMap(All, Invert_F{&NumPaths}, false);
  Levels[round].del();
  vertexMap(All, ClearVisited_F{&Visited}, false);
  vertexMap(Levels[round - 1], BC_Back_Vertex_F{&Visited, &Dependencies, &NumPaths}, false);
  for (long r = round - 2; r >= 0; r--) {
    edgeMap(G, Levels[r + 1], BC_Back_F{&Dependencies, &Visited}, false);
    Levels[r + 1].del();
    vertexMap(Levels[r], BC_Back_Vertex_F{&Visited, &Dependencies, &NumPaths}, false);
  }
  vertexMap(All, BC_Final_F{&Dependencies, &NumPaths}, false);
  Levels[0].del();
  All.del();
  double *out = static_cast<double *>(malloc(n * sizeof(double)));
  Dependencies.download(out, n);
  return out;
}

// ---- SSSP_BF with unit lengths (include/BellmanFordUnweighted.h:29-130) --
struct BF_F {
  DeviceArray<int32_t> *ShortestPathLen;
  DeviceArray<uint8_t> *Visited;  // padded to a multiple of 4 bytes (claimed in words)
};
template <>
struct edge_op_traits<BF_F> {
  static constexpr int op = TG_OP_BF;
  static tg_op_args args(const BF_F &f) {
    tg_op_args a{};
    a.dist = f.ShortestPathLen->data();
    a.visited = f.Visited->data();
    return a;
  }
};
struct BF_Vertex_F {
  DeviceArray<uint8_t> *Visited;
};
template <>
struct vertex_op_traits<BF_Vertex_F> {
  static constexpr int op = TG_VOP_CLEAR_VISITED;
  static tg_vmap_args args(const BF_Vertex_F &f) {
    tg_vmap_args a{};
    a.visited = f.Visited->data();
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
  VertexSubset frontier(G, static_cast<uint32_t>(start), n);
  uint64_t round = 0;
  while (frontier.not_empty()) {
    if (round == n) {  // negative weight cycle rule of the reference
      ShortestPathLen.fill(-inf);
      break;
    }
    // sparse only in the reference (threshold 1)
    VertexSubset output = edgeMap(G, frontier, BF_F{&ShortestPathLen, &Visited}, true, 1);
    vertexMap(output, BF_Vertex_F{&Visited}, false);
    frontier.del();
    frontier = output;
    round++;
  }
  frontier.del();
  int32_t *out = static_cast<int32_t *>(malloc(n * sizeof(int32_t)));
  ShortestPathLen.download(out, n);
  return out;
}

// ---- Triangle counting (include/TC.h:51-87) ------------------------------
// countF + count_common over every edge with s > d runs as one device pass.
inline uint64_t TC(Graph &G) {
  uint64_t c = 0;
  tg_check(tg_tc(G.handle(), &c));
  return c;
}

}  // namespace graphstore
