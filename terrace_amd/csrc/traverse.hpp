// Ligra-style traversal over the tiered graph (include/Map.cpp,
// include/VertexSubset.cpp): frontiers as a sparse queue and/or a dense
// bitmap, edgeMap with the reference's push/pull switch, and the five
// algorithms of include/BFS.h, Pagerank.h, Components.h, BC.h and TC.h.
#pragma once

#include <string>

#include "graph.hpp"

namespace tg {

// Frontier (VertexSubset): sparse queue and/or dense bitmap, or "all".
struct Frontier {
  uint32_t n = 0;
  bool all = false;
  bool has_queue = false;  // queue[0..count) valid
  bool has_bits = false;   // bits valid
  uint64_t count = 0;      // |frontier|
  DevBuf<uint32_t> queue;
  DevBuf<uint32_t> bits;
  uint64_t words() const { return (static_cast<uint64_t>(n) + 31) / 32; }
};

enum class Op : int { BFS = 1, CC = 2, BC_FWD = 3, BC_BACK = 4, PR = 5, TC = 6, BF = 7 };

// Device state an op reads and writes (pointers into device memory).
struct OpState {
  int32_t *parent = nullptr;    // BFS
  uint32_t *level = nullptr;    // BFS (optional)
  uint32_t depth = 0;           // BFS: level assigned to newly reached vertices
  uint32_t *ids = nullptr;      // CC
  const uint32_t *prev = nullptr;  // CC prevIDs (output = ids != prev)
  double *val = nullptr;        // BC NumPaths / Dependencies, PR p_next
  const double *src_val = nullptr;  // PR normalised p_curr
  const uint8_t *visited = nullptr;  // BC; BF claims it (writes), so it is 4-byte padded there
  int32_t *dist = nullptr;           // BF ShortestPathLen
};

// Sliced PageRank: the adjacency in 8 slices (one per XCD), the merge-path
// tiles of each slice and the offset of each slice's tiles.
constexpr uint32_t kPrsSlices = 8;
struct PrsTiles {
  uint32_t tiles[kPrsSlices];
  uint32_t base[kPrsSlices];
};

class Traversal {
 public:
  // s = null: use the graph's stream.  Traversals with their own streams
  // may run concurrently on one graph while no update is in flight.
  explicit Traversal(Graph &g, hipStream_t s = nullptr) : g_(g), stream_(s ? s : g.stream()) {}
  ~Traversal();
  Traversal(const Traversal &) = delete;
  Traversal &operator=(const Traversal &) = delete;
  hipStream_t stream() const { return stream_; }

  // Frontier helpers
  void frontier_single(Frontier &f, uint32_t v);
  void frontier_all(Frontier &f);
  void frontier_empty(Frontier &f);
  void to_bits(Frontier &f);
  void to_queue(Frontier &f);

  // edgeMap(G, in, F, output, threshold) (include/Map.cpp:66-74): dense
  // pull when n / threshold <= |in|, else sparse push.  out receives the
  // output frontier when output is true.
  void edge_map(Frontier &in, Op op, const OpState &st, Frontier *out, uint32_t threshold = 20);

  // vertexMap over the members of f (op codes of tg_vop).
  struct VArgs {
    double *val = nullptr;
    const double *inv = nullptr;
    const uint32_t *ids = nullptr;
    uint32_t *prev = nullptr;
    uint8_t *visited = nullptr;
  };
  void vertex_map(Frontier &f, int vop, const VArgs &a);
  void copy_frontier(Frontier &dst, Frontier &src);

  // Whole algorithms (device result buffers).  threshold 0 = the graph's
  // tuned value (Graph::tune).
  uint32_t bfs(uint32_t src, int32_t *parent, uint32_t *level, uint32_t threshold = 0);
  void pagerank(uint32_t iters, double *out);
  uint32_t cc(uint32_t *ids, uint32_t threshold = 0);
  void bc(uint32_t src, double *out, uint32_t threshold = 0);
  // SSSP_BF (include/BellmanFordUnweighted.h:60-130; on a weighted graph
  // include/BellmanFord.h:67-143 with the stored weights as lengths):
  // dist = shortest distance, INT_MAX / 2 unreachable; returns the rounds.
  uint32_t sssp_bf(uint32_t src, int32_t *dist, uint32_t threshold = 0);
  uint64_t tc();
  // 64 BFS in one traversal: levels[j*n + v] for source j (k <= 64).
  uint32_t bfs_multi(const uint32_t *srcs_host, uint32_t k, uint32_t *levels, uint32_t threshold = 0);
  // 64 BC in one traversal: out[j*n + v] for source j (k <= 64), each
  // equal to bc(srcs[j]) summed in sorted neighbour order.
  void bc_multi(const uint32_t *srcs_host, uint32_t k, double *out, uint32_t threshold = 0);
  // Triangles through each queried vertex (host in / host out).
  void ltc(const uint32_t *verts_host, uint32_t q, uint64_t *out_host);
  // Autotuner: sweeps the push/pull thresholds of BFS, multi-source BFS and
  // CC and the pull block size on the current graph and stores the fastest
  // in Graph::tune.  report[12]: thr_bfs, thr_ms, thr_cc, pull_block, then
  // (reference setting ms, tuned ms) for BFS, MS-BFS, CC and PageRank.
  void autotune(double *report);
  // Order-independent digest of a device array (u32 or u64 words).
  uint64_t digest(const uint32_t *a32, const uint64_t *a64, uint64_t n);

  // Per-kernel timing on the graph's stream (hipEvents), keyed by name.
  struct KTime {
    std::string name;
    double ms;
    uint64_t launches;
  };
  bool profiling = false;
  std::vector<KTime> ktimes;
  void record(const char *name, hipEvent_t a, hipEvent_t b);

  uint64_t last_edges_visited = 0;  // Σ degree over processed frontiers
  DevBuf<uint8_t> &tree_parts() { return tree_parts_; }
  const uint32_t *tree_task_vertex() const { return ttask_v_.get(); }
  const uint32_t *tree_task_block() const { return ttask_b_.get(); }

 private:
  Graph &g_;
  hipStream_t stream_;
  ScanScratch scan_;
  DevBuf<uint32_t> cnt_;        // device counters
  DevBuf<uint32_t> sec_pma_, sec_tree_;  // frontier vertices with secondary runs
  DevBuf<uint64_t> ttask_off_;  // tree task prefix over listed tree vertices
  DevBuf<uint32_t> ttask_v_, ttask_b_;  // tree task -> its vertex, its leaf block
  DevBuf<double> dtmp_, dtmp2_;
  DevBuf<uint32_t> utmp_;
  DevBuf<uint8_t> btmp_;
  DevBuf<uint64_t> csr_off_;
  DevBuf<uint32_t> csr_nbr_;
  DevBuf<uint64_t> red_;
  // per-call inputs and counters (sources, query vertices, their offsets
  // and counts), kept so that a call allocates nothing once warm: hipFree
  // waits for the whole device, so a call freeing a temporary would stall
  // work on every other stream
  DevBuf<uint32_t> qv_;
  DevBuf<uint64_t> qoff_, qcnt_;
  DevBuf<uint64_t> ms_vis_, ms_fr_, ms_nx_;
  DevBuf<uint64_t> bc_pm_, bc_mask_;  // multi-source BC: dense masks, per-level masks
  DevBuf<uint32_t> bc_list_;          // per-level vertex lists
  DevBuf<double> bc_sigma_, bc_dep_;  // [vertex][64]
  uint64_t ntree_tasks_ = 0;
  DevBuf<uint8_t> tree_parts_;  // per-task partials of ordered (FP64) tree pulls
  uint64_t tasks_epoch_ = ~0ull;  // graph epoch the tree tasks were built for
  const uint32_t *tree_list_ = nullptr;
  uint32_t ntree_list_ = 0;
  void prepare_tree_tasks();
  // PageRank over the XCD-sliced adjacency copy (graphs up to 6M vertices)
  // m: the graph's edge count (at most 2^32 - 1)
  void prepare_pr_slices(uint64_t m);
  void pagerank_sliced(uint32_t iters, double *out, uint64_t m);
  uint64_t prs_epoch_ = ~0ull, prs_m_ = 0;
  DevBuf<uint32_t> prs_off_, prs_nbr_, prs_deg_, prs_bnd_, prs_crow_, prs_coord_, prs_tc_;
  DevBuf<uint64_t> prs_meta_;
  DevBuf<double> prs_part_, prs_cval_;
  PrsTiles prs_tiles_{};
  uint32_t prs_max_tiles_ = 0, prs_ntiles_ = 0;
  void pull(Frontier &in, Op op, const OpState &st, Frontier *out);
  void push(Frontier &in, Op op, const OpState &st, Frontier *out);
  uint64_t popcount_bits(const Frontier &f);

  // A captured HIP graph and the launch parameters it was captured with.
  struct CachedGraph {
    std::vector<uint64_t> key;   // what exec was captured for
    std::vector<uint64_t> seen;  // the parameters of the last call
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void reset();
  };
  CachedGraph bfs_graph_, cc_graph_, bc_graph_, pr_graph_, prs_graph_;
  // The exec of cg, captured again from body (kernel launches on the
  // traversal's stream) when key differs from the one it was captured for
  // and equals the previous call's; null (the caller launches body's
  // kernels itself) on the first call with a key.
  template <class Body>
  hipGraphExec_t cached_graph(CachedGraph &cg, const std::vector<uint64_t> &key, const Body &body);

  // BFS / CC levels on device-resident counts: control words, two
  // frontiers (queue and bitmap each), the push lists.
  DevBuf<uint32_t> lvl_ctl_, lvl_q_[2], lvl_bits_[2], lvl_sec_pma_, lvl_sec_tree_;
  // BC: the forward levels' vertex lists and offsets, the backward frontier
  DevBuf<uint32_t> bc_lv_;
  DevBuf<uint64_t> bc_loff_;
  Frontier bc_back_;
  void level_buffers();
  // bc_*: a BC forward level (its vertices listed into bc_lv / bc_loff and
  // marked visited)
  template <class PushOp, class PullOp>
  void run_level(uint32_t x, uint32_t y, uint32_t thr, const PushOp &push_op, const PullOp &pull_op,
                 bool fused_pull, uint32_t *bc_lv = nullptr, uint64_t *bc_loff = nullptr,
                 uint8_t *bc_visited = nullptr);
  // Runs two-level steps (0 -> 1 -> 0) until frontier 0 is empty, replayed
  // from a HIP graph captured for key when the graph tuning asks for it;
  // returns the levels that ran.
  template <class Step>
  uint32_t run_levels(CachedGraph &cg, const std::vector<uint64_t> &key, const Step &step);
};

// Frontiers not yet bound to a graph (the reference's VertexSubset(e, n,
// all) carries no graph): vertex ops that read no graph state, and copies,
// on stream s.
void vertex_map_unbound(Frontier &f, int vop, const Traversal::VArgs &a, hipStream_t s);
void copy_frontier_unbound(Frontier &dst, Frontier &src, hipStream_t s);

}  // namespace tg
