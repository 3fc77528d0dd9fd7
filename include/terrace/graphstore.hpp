// graphstore.hpp — C++ drop-in for the reference's graphstore::Graph,
// VertexSubset, edgeMap and vertexMap (include/graph.h, include/Map.cpp,
// include/VertexSubset.cpp), implemented over the C ABI of
// libterrace_hip.so (terrace_hip.h).  Header-only; link with
// -lterrace_hip.
//
// Algorithm code keeps the reference's shape: functor structs passed to
// edgeMap / vertexMap, built with the reference's constructor arguments.
// Device functors cannot cross a C ABI, so each functor type maps to a
// device op through edge_op_traits / vertex_op_traits; a functor without a
// trait is rejected at compile time.  Functor state lives in HBM
// (DeviceArray<T>, which converts to the T* the functors hold), not in
// host arrays.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../terrace_hip.h"

namespace graphstore {

inline void tg_check(int rc) {
  if (rc != TG_OK) throw std::runtime_error(std::string("terrace_hip: ") + tg_last_error());
}

// Device array owned by host code (functor state).
template <typename T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(uint64_t n) : n_(n) { tg_check(tg_dev_alloc(n * sizeof(T), reinterpret_cast<void **>(&p_))); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  ~DeviceArray() {
    if (p_) tg_dev_free(p_);
  }
  T *data() const { return p_; }
  operator T *() const { return p_; }  // functors hold the device pointer
  uint64_t size() const { return n_; }
  void upload(const T *h, uint64_t count) { tg_check(tg_memcpy_h2d(p_, h, count * sizeof(T))); }
  void download(T *h, uint64_t count) const { tg_check(tg_memcpy_d2h(h, p_, count * sizeof(T))); }
  void fill(const T &v) {
    std::unique_ptr<T[]> h(new T[n_ ? n_ : 1]);
    std::fill(h.get(), h.get() + n_, v);
    upload(h.get(), n_);
  }
  void set(uint64_t i, const T &v) { tg_check(tg_memcpy_h2d(p_ + i, &v, sizeof(T))); }
  std::unique_ptr<T[]> to_host() const {
    std::unique_ptr<T[]> h(new T[n_ ? n_ : 1]);
    download(h.get(), n_);
    return h;
  }

 private:
  T *p_ = nullptr;
  uint64_t n_ = 0;
};

// graphstore::Graph (include/graph.h:73-230), resident in HBM.
class Graph {
 public:
  typedef uint32_t vertex;
  typedef uint32_t weight;
  typedef std::pair<vertex, vertex> edge;

  // NeighborIterator (include/graph.h:93-116): sorted neighbour stream.
  class NeighborIterator {
   public:
    NeighborIterator(const Graph *g, vertex v) : list_(g->neighbor_list(v)) {}
    vertex operator*() const { return list_[i_]; }
    void operator++() { ++i_; }
    bool done() const { return i_ >= list_.size(); }

   private:
    std::vector<vertex> list_;
    size_t i_ = 0;
  };

  explicit Graph(uint32_t size) : n_(size) { tg_check(tg_create(size, &h_)); }
  // The WEIGHTED build's graph (include/graph.h:34-38) as a run-time choice.
  Graph(uint32_t size, bool weighted) : n_(size) {
    tg_check(weighted ? tg_create_weighted(size, &h_) : tg_create(size, &h_));
  }
  ~Graph() {
    if (h_) tg_destroy(h_);
  }
  Graph(const Graph &) = delete;
  Graph &operator=(const Graph &) = delete;

  // graph.h:507-641 / :675-791: 0 on success, -1 if present / absent.
  int add_edge(const vertex s, const vertex d) {
    int r = 0;
    tg_check(tg_add_edge(h_, s, d, &r));
    return r;
  }
  int remove_edge(const vertex s, const vertex d) {
    int r = 0;
    tg_check(tg_remove_edge(h_, s, d, &r));
    return r;
  }
  // graph.h:434-502.  The inputs are not modified and need not be sorted;
  // perm is accepted for signature compatibility.
  void add_edge_batch(vertex *srcs, vertex *dests, uint32_t edge_count, std::vector<uint32_t> &perm) {
    (void)perm;
    tg_check(tg_insert_batch(h_, srcs, dests, edge_count));
  }
  // graph.h:430-436 (WEIGHTED): the first weight of an edge in the batch
  // counts; existing edges keep their weight.
  void add_edge_batch(vertex *srcs, vertex *dests, weight *wghts, uint32_t edge_count, std::vector<uint32_t> &perm) {
    (void)perm;
    tg_check(tg_insert_batch_weighted(h_, srcs, dests, wghts, edge_count));
  }
  // is_edge of the weighted build returns the weight (0xFFFFFFFF here when
  // absent).
  weight edge_weight(const vertex s, const vertex d) {
    weight w = 0;
    tg_check(tg_edge_weights(h_, &s, &d, 1, &w));
    return w;
  }
  void remove_edge_batch(const vertex *srcs, const vertex *dests, uint32_t edge_count) {
    tg_check(tg_delete_batch(h_, srcs, dests, edge_count));
  }
  uint32_t is_edge(const vertex s, const vertex d) {
    uint32_t r = 0;
    tg_check(tg_is_edge(h_, s, d, &r));
    return r;
  }
  NeighborIterator neighbors(const vertex v) const { return NeighborIterator(this, v); }
  uint32_t degree(const vertex v) const {
    uint32_t r = 0;
    tg_check(tg_degree(h_, v, &r));
    return r;
  }
  uint64_t get_size(void) {
    uint64_t r = 0;
    tg_check(tg_memory_bytes(h_, &r));
    return r;
  }
  // graph.h:817-822 returns uint32; the full count is num_edges64().
  uint32_t get_num_edges(void) { return static_cast<uint32_t>(num_edges64()); }
  uint64_t num_edges64(void) {
    uint64_t r = 0;
    tg_check(tg_num_edges(h_, &r));
    return r;
  }
  uint32_t get_num_vertices(void) const { return n_; }
  // graph.h:1085-1255: common neighbours of s and d below min(s, d), on
  // the device (tg_count_common).
  uint32_t count_common(const vertex s, const vertex d) const {
    uint32_t c = 0;
    tg_check(tg_count_common(h_, &s, &d, 1, &c));
    return c;
  }
  // Many pairs in one call: out[i] = count_common(srcs[i], dests[i]).
  void count_common(const vertex *srcs, const vertex *dests, uint64_t q, uint32_t *out) const {
    tg_check(tg_count_common(h_, srcs, dests, q, out));
  }
  std::vector<vertex> neighbor_list(vertex v) const {
    uint32_t deg = 0;
    tg_check(tg_neighbors(h_, v, nullptr, 0, &deg));
    std::vector<vertex> out(deg);
    if (deg) tg_check(tg_neighbors(h_, v, out.data(), deg, &deg));
    return out;
  }
  tg_graph *handle() const { return h_; }

 private:
  tg_graph *h_ = nullptr;
  uint32_t n_;
};

// VertexSubset (include/VertexSubset.cpp).  VertexSubset(e, n) holds one
// vertex, VertexSubset(0, n, true) every vertex, VertexSubset(els, n) the
// vertices whose els[i] is true; none of them names a graph, as in the
// reference: the first edgeMap binds the subset to its graph.  Copies share
// the device frontier as the reference's shallow copies do; del() releases
// it.
class VertexSubset {
 public:
  VertexSubset() = default;
  // VertexSubset(el_t e, uint64_t max_el, bool all) (VertexSubset.cpp:109)
  VertexSubset(uint32_t e, uint64_t max_el, bool all = false) : n_(max_el) {
    tg_check(tg_frontier_create_n(static_cast<uint32_t>(max_el), &f_));
    if (all) tg_check(tg_frontier_set_all(f_));
    else tg_check(tg_frontier_set(f_, &e, 1));
  }
  // VertexSubset(bool *els, uint64_t len) (VertexSubset.cpp:119)
  VertexSubset(bool *els, uint64_t len) : n_(len) {
    static_assert(sizeof(bool) == 1, "dense subsets are passed as bytes");
    tg_check(tg_frontier_create_n(static_cast<uint32_t>(len), &f_));
    tg_check(tg_frontier_set_dense(f_, reinterpret_cast<const uint8_t *>(els), len));
  }
  // The same bound to g from the start.
  VertexSubset(Graph &g, uint32_t e, uint64_t max_el, bool all = false) : n_(max_el) {
    tg_check(tg_frontier_create(g.handle(), &f_));
    if (all) tg_check(tg_frontier_set_all(f_));
    else tg_check(tg_frontier_set(f_, &e, 1));
  }
  static VertexSubset empty(uint64_t n) {
    VertexSubset v;
    v.n_ = n;
    tg_check(tg_frontier_create_n(static_cast<uint32_t>(n), &v.f_));
    return v;
  }
  uint64_t get_n() const {
    uint64_t r = 0;
    tg_check(tg_frontier_size(f_, &r));
    return r;
  }
  bool not_empty() const { return get_n() > 0; }
  std::vector<uint32_t> members() const {
    uint64_t n = get_n(), c = 0;
    std::vector<uint32_t> out(n);
    tg_check(tg_frontier_get(f_, out.data(), n, &c));
    out.resize(c);
    return out;
  }
  VertexSubset copy() const {
    VertexSubset v = empty(n_);
    tg_check(tg_frontier_copy(v.f_, f_));
    return v;
  }
  void del() {
    if (f_) tg_frontier_destroy(f_);
    f_ = nullptr;
  }
  tg_frontier *handle() const { return f_; }
  uint64_t max_el() const { return n_; }

 private:
  uint64_t n_ = 0;
  tg_frontier *f_ = nullptr;
};

// Trait tables: functor type -> device op and its arguments.
template <class F, class = void>
struct edge_op_traits;  // specialised next to each functor
template <class F, class = void>
struct vertex_op_traits;

template <class F, class = void>
struct has_edge_op : std::false_type {};
template <class F>
struct has_edge_op<F, std::void_t<decltype(edge_op_traits<F>::op)>> : std::true_type {};
template <class F, class = void>
struct has_vertex_op : std::false_type {};
template <class F>
struct has_vertex_op<F, std::void_t<decltype(vertex_op_traits<F>::op)>> : std::true_type {};

// The graph a vertex functor holds (PR_Vertex(p, G) does), else none.
template <class F, class = void>
struct functor_graph {
  static tg_graph *get(const F &) { return nullptr; }
};
template <class F>
struct functor_graph<F, std::void_t<decltype(std::declval<const F &>().G.handle())>> {
  static tg_graph *get(const F &f) { return f.G.handle(); }
};

// edgeMap (include/Map.cpp:66-74).
template <class F>
VertexSubset edgeMap(Graph &G, VertexSubset &vs, F f, bool output = true, uint32_t threshold = 20) {
  static_assert(has_edge_op<F>::value, "edgeMap: functor has no device op (edge_op_traits)");
  tg_op_args a = edge_op_traits<F>::args(f);
  VertexSubset out = output ? VertexSubset::empty(G.get_num_vertices()) : VertexSubset();
  tg_check(tg_edge_map(G.handle(), vs.handle(), edge_op_traits<F>::op, &a, output ? out.handle() : nullptr,
                       threshold));
  return out;
}

// vertexMap (include/Map.cpp:76-108).  Every vertex functor of the
// algorithm headers returns 1, so the output subset is a copy of vs.
template <class F>
VertexSubset vertexMap(VertexSubset &vs, F f, bool output = true) {
  static_assert(has_vertex_op<F>::value, "vertexMap: functor has no device op (vertex_op_traits)");
  tg_vmap_args a = vertex_op_traits<F>::args(f);
  tg_check(tg_vertex_map(functor_graph<F>::get(f), vs.handle(), vertex_op_traits<F>::op, &a));
  return output ? vs.copy() : VertexSubset();
}

}  // namespace graphstore
