// Degree-tiered dynamic graph resident in HBM (gfx950).
//
// Tiers, following include/graph.h:61-72 with the thresholds the code
// actually uses (graph.h:34-39):
//   * inline  : 13 smallest neighbours in a 64-byte vertex block
//   * PMA     : the rest of every vertex with degree <= 1024, in one shared
//               packed memory array sorted by (src,dst) with empty slots
//   * B-tree  : the rest of every vertex routed past 1024, as sorted leaf
//               blocks of 256 ids under a per-vertex index (a height-2
//               B+-tree whose root is the index array)
#pragma once

#include "prims.hpp"
#include "weights.hpp"

#include <memory>
#include <mutex>
#include <vector>

namespace tg {

// Raw device view handed to kernels.
struct GView {
  VBlock *vb;
  uint32_t n;
  // PMA
  uint32_t *pdst;
  uint32_t *psrc;
  uint64_t N;       // slots = nleaves * 64
  uint32_t H;       // nleaves = 1 << H
  uint32_t *pcnt;   // count tree, level l at tree_off(l)
  uint64_t *fence;  // [nleaves] lower key fence of each leaf
  // B-tree tier
  uint32_t *bkeys;  // [blocks * 256]
  uint32_t *bcnt;   // [blocks]
  uint64_t *toff;   // [trees] index segment offset
  uint32_t *tnblk;  // [trees] blocks in the index
  uint32_t *iblk;   // index pool: block id
  uint32_t *imin;   // index pool: first key of the block
};

__host__ __device__ inline uint64_t tree_level_off(uint32_t H, uint32_t l) {
  uint64_t nl = 1ull << H;
  return 2 * nl - ((2 * nl) >> l);
}

// Words of a graph's batch scalars: one device array of u64 that a batch
// reads back in one copy.
//   [0] error bits; [1..13] the counts and in-place statuses of one batch
//   (insert half: [3] PMA keys, [4] tree keys, [5] promotions, [7] PMA
//   leaf-fit flag, [8] touched leaves, [13] in-place tree status);
//   [14], [15] look-back error words (sort, scans/selects);
//   [16..19] the inline refill's tier split and tree status;
//   [20] a PMA delete left a leaf at or below the leaf lower bound;
//   [21..23] the insert half's gate: the key and run counts its kernels
//   use (zero when it does not run) and 1 ran / 2 did not run;
//   [24..47] the delete half's words, at the same offsets from kDel;
//   [48], [49] key and run counts the host writes for a half it launches
//   on counts it knows.
// A mixed batch launches both halves before any read, so the delete half
// keeps its counts apart from the insert half's.
namespace bsw {
constexpr int kErrSort = 14, kErrScan = 15;
constexpr int kLowFlag = 20;
constexpr int kGate = 21;  // [kGate] keys, [kGate + 1] runs, [kGate + 2] 1 ran / 2 did not
constexpr int kDel = 24;   // the delete half's block
constexpr int kPut = 48;
constexpr int kCount = 56;
}  // namespace bsw

struct Bounds {
  uint64_t up[48];  // max elements a window at level l may hold
  uint64_t lo[48];  // min elements a window at level l should hold
};

struct Stats {
  uint64_t pma_rebuilds = 0;
  uint64_t pma_windows = 0;
  uint64_t pma_window_leaves = 0;
  uint64_t tree_blocks_written = 0;
  uint64_t promotions = 0;
  uint64_t tree_small_updates = 0;  // tree updates done in place
  uint64_t pma_leaf_merges = 0;     // PMA inserts merged inside their leaves
  uint64_t pma_low_windows = 0;     // underfull PMA windows respread after deletes
  uint64_t pma_shrinks = 0;         // PMA arrays shrunk after deletes
  // the last batch: unique inserted keys and their sources, existing
  // deleted keys and their sources
  uint64_t last_ins_keys = 0, last_ins_sources = 0, last_del_keys = 0, last_del_sources = 0;
  // over every update batch since the graph was created
  uint64_t sum_ins_sources = 0, sum_del_sources = 0;
};

class Graph {
 public:
  explicit Graph(uint32_t n);
  ~Graph();
  Graph(const Graph &) = delete;
  Graph &operator=(const Graph &) = delete;

  uint32_t num_vertices() const { return n_; }
  hipStream_t stream() const { return stream_; }
  GView view() const;

  // Batch updates.  src/dst are device pointers of count ids each; the
  // inputs are not modified.  Any order, duplicates allowed.
  void insert_batch(const uint32_t *src, const uint32_t *dst, uint64_t count);
  void delete_batch(const uint32_t *src, const uint32_t *dst, uint64_t count);
  // A mixed batch (device pointers): the ni inserts, then the nd deletes,
  // with one sort and preparation for both; equals insert_batch(is, id, ni)
  // followed by delete_batch(ds, dd, nd).
  void update_batch(const uint32_t *is, const uint32_t *id, uint64_t ni, const uint32_t *ds, const uint32_t *dd,
                    uint64_t nd);
  // Sorted unique keys (src<<32|dst) already in HBM.
  void insert_sorted(const uint64_t *keys, uint64_t n);
  void delete_sorted(const uint64_t *keys, uint64_t n);
  // Coalesced update batches: ops[i] 0 = insert, 1 = delete, arrival order;
  // the result equals applying them one after another.
  void apply_ops(const uint32_t *src, const uint32_t *dst, const uint32_t *ops, uint64_t count);

  // Weighted graphs (the reference's WEIGHTED build): one uint32 weight per
  // directed edge, below 2^31 (the Bellman-Ford edge length is an int32).
  // A batch keeps the first weight given for a key; inserting an edge that
  // exists leaves its weight unchanged; deleting an edge drops its weight.
  void enable_weights();
  bool weighted() const { return ws_ != nullptr; }
  void insert_batch_weighted(const uint32_t *src, const uint32_t *dst, const uint32_t *w, uint64_t count);
  WView weight_view() const { return ws_ ? ws_->view() : WView{}; }
  // out[i] = weight of (src[i], dst[i]), or 0xFFFFFFFF when absent
  void edge_weights(const uint32_t *src, const uint32_t *dst, uint64_t count, uint32_t *out);

  // Queries (device pointers).
  void degrees(uint32_t *out) const;
  uint64_t num_edges();
  void is_edge_batch(const uint32_t *src, const uint32_t *dst, uint64_t count, uint8_t *out);
  // Graph::count_common (include/graph.h:901-1255) for q pairs (device
  // arrays): out[i] = common neighbours of src[i] and dst[i] below both.
  // err: a device word for the id check (null: the graph's own, allocated
  // on first use).
  void count_common(const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out, uint64_t *err = nullptr);
  // TC.h's edgeMap(G, F, countF) (include/TC.h:51-87): adds to *total (a
  // device counter) the sum of count_common(u, v) over the frontier's
  // vertices u (list of cnt, or every vertex when list is null) and their
  // neighbours v < u.  Returns the number of such edges.
  uint64_t count_common_edges(const uint32_t *list, uint64_t cnt, unsigned long long *total);
  // offsets[n+1] (device, u64) and neighbours (device, u32), sorted per vertex.
  void export_adjacency(uint64_t *offsets, uint32_t *nbrs);
  // Sorted neighbours of one vertex into out (device, >= degree entries).
  void neighbors_of(uint32_t v, uint32_t *out);
  // 0 = inline only, 1 = PMA, 2 = B-tree, per vertex (device u8[n]).
  void tiers(uint8_t *out) const;
  // vertices and secondary edges per tier: PMA count, PMA edges, tree count, tree edges
  void tier_counts(uint64_t out[4]);
  uint64_t memory_bytes() const;
  uint32_t num_trees() const { return ntrees_; }
  uint64_t pma_slots() const { return pma_N_; }
  uint64_t pma_elements() const { return pma_elems_; }
  // an upper bound on the edges, known on the host: the inline slots, the
  // PMA's elements and the capacity of the tree blocks in use
  uint64_t edge_bound() const { return 13ull * n_ + pma_elems_ + 256ull * blk_top_; }
  // leaf counts into host memory (at most cap); returns the number of leaves
  uint64_t pma_leaf_counts(uint32_t *host, uint64_t cap);
  const Stats &stats() const { return stats_; }
  // True once a batch failed after it had begun changing the graph (an
  // allocation for the graph's own growth, or an internal check): the
  // adjacency may then be half-updated, so every later call on the graph
  // fails with code -4.  A batch that fails before it changes anything
  // (bad ids, a scratch allocation, a preparation check) leaves the graph
  // valid and unchanged.
  bool invalid() const { return invalid_; }
  // Failure injection for tests of the error paths (every kernel still runs
  // to its normal end): bit 0 makes the next batch's preparation report a
  // look-back error word; bit 1 makes the next batch fail as if an
  // allocation failed right after its first write to the graph.
  uint32_t debug_fail = 0;

  // Versioned snapshot: make this graph a read-only copy of src's current
  // state (device-to-device copies ordered on src's stream).  Queries can
  // then run on this copy while src takes the next update batch.
  // query_only: skip what only updates read (PMA sources, count tree,
  // fences); such a snapshot rejects updates and existence checks.
  void copy_from(Graph &src, bool query_only = false);

  // Vertex lists by tier, rebuilt lazily after updates (device arrays).
  struct TierLists {
    uint32_t *pma = nullptr;   // vertices with a PMA run
    uint32_t npma = 0;
    uint32_t *tree = nullptr;  // vertices with a B-tree
    uint32_t ntree = 0;
  };
  const TierLists &tier_lists();

  // Largest tree update rewritten in place (see tree_small).
  static constexpr uint64_t kTreeSmallKeys = 4096;

  // Mutation counter: changes whenever the adjacency may have changed.
  uint64_t epoch() const { return epoch_; }
  // Drops what was derived from the current version (the tier lists, and
  // every traversal cache keyed by the epoch) without changing the
  // adjacency: the next query rebuilds them as after an update.
  void touch() {
    lists_dirty_ = true;
    ++epoch_;
  }

  // Launch parameters chosen by the autotuner; the whole-algorithm calls
  // use them.  thr_*: edgeMap push/pull switch (dense when n / thr <=
  // |frontier|, include/Map.cpp:66-74, reference value 20; BFS starts at
  // 40 and CC at 5, the fastest in threshold sweeps on the scale-22 plain
  // and skewed rMAT graphs); pull_block: workgroup size of the inline and
  // PMA pull kernels.  Results do not depend on them.
  // pr_graph: PageRank iterations replayed from a captured HIP graph.
  // pr_sliced: PageRank over the XCD-sliced adjacency copy when the graph
  // has at most 6M vertices (else the tiered pull).
  // push_wave: push frontiers up to this size give each PMA vertex a wave
  // (64 lanes), larger ones a 16-lane group.
  // level_graph: BFS and CC levels (directions chosen on the device)
  // replayed from a captured HIP graph of two levels; 0 launches the same
  // kernels one by one.
  struct Tuning {
    uint32_t thr_bfs = 40, thr_ms = 20, thr_cc = 5;
    uint32_t pull_block = 256;
    uint32_t pr_graph = 1;
    uint32_t pr_sliced = 1;
    uint32_t push_wave = 16384;
    uint32_t level_graph = 1;
  };
  Tuning tune;

  std::mutex mu;  // serialises calls on one handle
  ScanScratch scan_scratch;

 private:
  uint32_t n_;
  hipStream_t stream_ = nullptr;
  // A second stream for the B-tree part of a large batch, which touches
  // no PMA data: it runs beside the PMA kernels, forked from and joined
  // back to stream_ by events (every fork is joined before the batch's
  // read, so synchronising stream_ covers it).
  hipStream_t side_ = nullptr;
  hipEvent_t ev_[4] = {};
  static constexpr uint64_t kSideMin = 1u << 14;  // batches from this many keys use it
  void fork_side(int e);  // side_ waits for stream_'s work so far
  void join_side(int e);  // stream_ waits for side_'s work so far
  DevBuf<VBlock> vb_;
  // PMA
  uint64_t pma_N_ = 0;
  uint32_t pma_H_ = 0;
  DevBuf<uint32_t> pdst_, psrc_, pcnt_;
  DevBuf<uint64_t> fence_;
  DevBuf<uint8_t> mark_;
  Bounds bounds_{};
  bool tree_dirty_ = false;  // count tree levels above the leaves are stale
  uint64_t pma_elems_ = 0;   // elements in the PMA (the count tree's root, kept on the host)
  bool read_only_ = false;   // query-only snapshot
  bool invalid_ = false;     // see invalid()
  // B-tree tier
  DevBuf<uint32_t> bkeys_, bcnt_;
  uint64_t blk_top_ = 0;
  DevBuf<uint32_t> iblk_, imin_;
  uint64_t idx_top_ = 0;
  DevBuf<uint64_t> toff_;
  DevBuf<uint32_t> tnblk_, tvtx_;
  uint32_t ntrees_ = 0;
  Stats stats_;

  // tier lists
  bool lists_dirty_ = true;
  uint64_t epoch_ = 0;
  std::unique_ptr<WeightStore> ws_;  // edge weights (weighted graphs only)
  DevBuf<uint32_t> wv0_, wv1_, wput_;
  TierLists lists_;
  DevBuf<uint32_t> list_pma_, list_tree_;

  // scratch
  SortScratch sort_;
  DevBuf<uint64_t> k0_, k1_, k2_, k3_;
  DevBuf<uint64_t> kD_, kE_;  // mixed batches: deleted edges, the existing ones among them
  DevBuf<uint32_t> fD_;       // mixed batches: existence flags of the deleted edges
  DevBuf<uint32_t> f0_, f1_, f2_, f3_;
  DevBuf<uint64_t> g0_, g1_, g2_, g3_;
  DevBuf<uint32_t> u0_, u1_, u2_, u3_, u4_;
  DevBuf<uint32_t> kick_;
  DevBuf<uint8_t> route_;
  DevBuf<uint64_t> scal_;  // small device scalars
  // batch scalars (word layout: namespace bsw)
  static constexpr int kScalars = bsw::kCount;
  static constexpr int kErrSort = bsw::kErrSort, kErrScan = bsw::kErrScan;
  static constexpr int kLowFlag = bsw::kLowFlag;
  DevBuf<uint64_t> dsc_;
  DevBuf<uint64_t> cc_err_, cc_off_;  // count_common: id check; CSR offsets of the TC edgeMap
  DevBuf<uint32_t> cc_nbr_;
  DevBuf<uint64_t> tmp_old_;
  DevBuf<uint32_t> ops_val_;
  DevBuf<uint64_t> ops_ins_, ops_del_;
  // per-batch scratch kept across calls (no allocation on the update path
  // once the buffers have grown)
  DevBuf<uint32_t> s32_[32];
  DevBuf<uint8_t> s8_[8];
  DevBuf<uint64_t> s64_[4];
  // mixed batches: the PMA slot of each deleted key as the preparation
  // found it, and the same hints for the PMA delete list
  DevBuf<uint64_t> dslot_, dhint_;
  // bucket sort of mid-size batches: per-bucket counts and cursors (zero
  // between calls), bucket starts, the bucketed keys
  DevBuf<uint32_t> sbcnt_, sbcur_, sboff_;
  DevBuf<uint64_t> sbkeys_;

  // Per-batch scratch above kKeepScratch elements is released when a batch
  // that large ends (a static build of 1.5B edges would otherwise hold tens
  // of GB of sort and flag buffers for the graph's life); smaller buffers
  // stay, so steady-state batches allocate nothing.
  static constexpr uint64_t kKeepScratch = 1ull << 25;
  void trim_scratch() noexcept;
  struct TrimAfter {  // releases the scratch when a batch of `count` ends
    Graph &g;
    bool on;
    ~TrimAfter() {
      if (on) g.trim_scratch();
    }
  };

  void init_pma(uint64_t nleaves);
  void compute_bounds();
  // true when the bucket sort ran (its overflow bit may then be set)
  bool prepare_keys_async(const uint32_t *src, const uint32_t *dst, uint64_t count, bool radix = false);
  bool prepare_mixed_async(const uint32_t *is, const uint32_t *id, uint64_t ni, const uint32_t *ds,
                           const uint32_t *dd, uint64_t nd, bool radix = false);
  // presize, then launch + read + finish under mutate().  cw: the key and
  // run counts on the device (two adjacent words), at most nR and nruns.
  void insert_runs(const uint64_t *R, uint64_t nR, uint32_t nruns, const uint64_t *cw);
  // run_beg: the runs' starts (null: s32_[0])
  void delete_runs(const uint64_t *R, uint64_t nR, uint32_t nruns, const uint64_t *cw,
                   const uint32_t *run_beg = nullptr);
  // The insert half's kernels, sized by capacities (capR keys, capRuns
  // runs), on the counts in cw; gate (the batch scalars, or null) closes
  // the half when the preparation raised an error; ileaf (optional): the
  // PMA leaf each key's existence lookup searched (~0: none).  Returns
  // whether the in-leaf PMA merge was launched.
  bool insert_launch(const uint64_t *R, uint64_t capR, uint32_t capRuns, const uint64_t *cw, const uint64_t *gate,
                     const uint32_t *ileaf = nullptr);
  // After the read: the insert's outcome, and its general passes when the
  // in-place paths did not take every key.  True when the in-place paths
  // took them all (flags not yet checked).
  bool insert_finish(const uint64_t *hs, const uint64_t *R, bool pma_launched);
  // The delete half's kernels, its words at ds.  spec: launched behind the
  // insert half of the same batch, it runs only when that half completed
  // on the device (pma_launched: the insert's in-leaf merge was launched).
  // R: the deleted keys, with exf their existence flags and dnidx the prefix
  // of the absent ones (both null: every key of R exists).  dslot (optional):
  // per key of R the PMA slot the preparation found it in (a hint).
  void delete_launch(const uint64_t *R, const uint32_t *exf, const uint32_t *dnidx, uint64_t capR,
                     uint32_t capRuns, const uint32_t *run_beg, const uint64_t *cw, const uint64_t *gate, bool spec,
                     bool pma_launched, uint64_t *ds, const uint64_t *dslot = nullptr);
  void delete_finish(const uint64_t *hs, const uint64_t *R, const uint32_t *run_beg, uint64_t *ds);
  // k_bump: the inline refill (tree keys it moves into s64_[3], count ds[17])
  void delete_refill(const uint64_t *R, uint32_t capRuns, const uint32_t *run_beg, uint64_t *ds, bool trees_only);
  // keys and runs written to the device (bsw::kPut); returns their address
  const uint64_t *put_counts(uint64_t keys, uint64_t runs);
  uint64_t compact_keys(const uint64_t *in, const uint32_t *flags, uint64_t n, uint64_t *out);
  void check_batch_flags();
  void check_flags(const uint64_t *h);
  // Error words of the preparation (id range, look-back waits) read back
  // before the first write to the graph: throws, leaving it unchanged.
  void check_prep(const uint64_t *h);
  // Runs body, which changes the graph; any failure marks the graph invalid.
  template <class F>
  void mutate(F &&body);
  // Per-batch scratch sized before the first write (so that running out of
  // memory there leaves the graph unchanged).
  void presize_insert(uint64_t nR, uint32_t nruns);
  void presize_delete(uint64_t nR, uint32_t nruns);
  // PMA inserts: pma_insert = in-leaf merge when every touched leaf fits,
  // else the window pass (or a rebuild beyond the root bound)
  void pma_insert(const uint64_t *keys, uint64_t k);
  // kdev: the key count on the device, at most cap (null: cap keys);
  // leaves_known: u0_ holds each key's leaf or ~0 where it must be searched
  // (k_insert_runs wrote it)
  void pma_fast_launch(const uint64_t *keys, uint64_t cap, const uint64_t *kdev = nullptr,
                       bool leaves_known = false);
  // hint (optional): per key a slot that may still hold it
  void pma_delete_launch(const uint64_t *keys, uint64_t cap, const uint64_t *kdev, const uint64_t *hint = nullptr);
  void pma_fast_done(uint64_t k, uint64_t nt);
  void pma_window_insert(const uint64_t *keys, uint64_t k);
  void pma_respread(const uint64_t *keys, uint64_t k, uint64_t cap, bool low);
  void pma_rebalance_low();
  void pma_rebuild(const uint64_t *keys, uint64_t k, uint64_t total);
  void rebuild_count_tree();
  void tree_insert(const uint64_t *keys, uint64_t k);
  void tree_delete(const uint64_t *keys, uint64_t k);
  // in-place update of up to 4096 tree keys when every touched block keeps
  // between 1 and 256 keys; false (nothing written) otherwise
  bool tree_small(const uint64_t *keys, uint64_t k, bool insert);
  // status (null: dsc_[13]): 0 done, 1 did not fit, 2 gate set
  void tree_small_launch(const uint64_t *keys, uint64_t cap, bool insert, const uint64_t *kdev = nullptr,
                         const uint64_t *gate = nullptr, uint64_t *status = nullptr, hipStream_t s = nullptr);
  void tree_insert_cow(const uint64_t *keys, uint64_t k);
  void tree_delete_cow(const uint64_t *keys, uint64_t k);
  void ensure_blocks(uint64_t need);
  void ensure_index(uint64_t need);
  void ensure_trees(uint64_t need);
};

}  // namespace tg
