// Request loop over one resident graph: interleaved update batches and
// queries (BFS, CC, PageRank, local triangle count) replayed at their
// arrival times.  No counterpart in the reference.
//
// plain   : one request at a time, in arrival order, through the same
//           calls a single user makes (insert_batch + delete_batch, bfs,
//           cc, pagerank, ltc).
// strict  : continuous batching.  Pending requests are cut into maximal
//           arrival-ordered runs of updates or of queries, so every query
//           sees exactly the updates that arrived before it (the version
//           it runs on).  A run of updates is coalesced into one device
//           pass (last op per edge wins, equal to applying them in order).
//           A run of queries is admitted while its memory footprint fits
//           the budget; BFS queries share multi-source traversals (64 per
//           lane mask), identical CC / PR queries share one computation,
//           LTC queries share one launch, and the query kinds run
//           concurrently on separate streams.
// admission: continuous batching with snapshot-at-admission.  Updates are
//           applied in arrival order as soon as they arrive (coalesced);
//           queries wait in a pending pool and are admitted together, as
//           slots and memory free up, on a snapshot of the version current
//           at admission.  Every query sees at least the updates that
//           arrived before it, and possibly updates that arrived while it
//           waited (a linearizable order: those updates and the query were
//           in flight together).  Each result carries its version and
//           equals the query run alone on that version.
#pragma once

#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#include "traverse.hpp"

namespace tg {

enum ReqType { kReqUpdate = 1, kReqBfs = 2, kReqCc = 3, kReqPr = 4, kReqLtc = 5 };
enum ServeMode { kServePlain = 0, kServeStrict = 1, kServeAdmission = 2 };

struct Request {
  int type = 0;
  double arrival = 0;
  uint32_t src = 0;
  uint32_t iters = 10;
  std::vector<uint32_t> us, ud, uop;  // update: (src, dst, 0 insert / 1 delete), in order
  // results
  int status = 1;  // 1 pending, 0 done, < 0 error code
  double start = 0, done = 0;
  uint64_t version = 0;
  uint64_t digest = 0;  // order-independent digest of the result array
  uint64_t value = 0;   // LTC count; BFS vertices reached; CC components
  double fvalue = 0;    // PR: sum of ranks
  uint64_t group = 0;   // id of the batch it ran in
  uint64_t footprint = 0;
  std::vector<uint32_t> r32;  // kept results (levels / labels)
  std::vector<double> r64;    // kept results (ranks)
};

// A versioned snapshot and the traversal lanes that query it.
struct QuerySlot {
  std::unique_ptr<Graph> snap;
  static constexpr int kLanes = 4;  // BFS, CC, PR, LTC
  hipStream_t streams[kLanes] = {};
  std::unique_ptr<Traversal> trav[kLanes];
  DevBuf<uint32_t> res32[kLanes];
  DevBuf<int32_t> par;
  DevBuf<double> res64;
  std::thread worker;
  std::atomic<bool> running{false};
  std::string error;
};

class Server {
 public:
  Server(Graph &g, int mode, uint64_t mem_budget, bool keep_results);
  ~Server();
  uint64_t submit(Request r);
  void run();
  const Request &result(uint64_t id) const { return reqs_.at(id); }
  // [0] makespan s, [1] requests/s, [2] p50 s, [3] p99 s, [4] p50 update,
  // [5] p99 update, [6] p50 query, [7] p99 query, [8] groups, [9] rejected
  std::vector<double> summary() const;

 private:
  Graph &g_;
  int mode_;
  bool batched_;
  uint64_t budget_;
  bool keep_;
  std::vector<Request> reqs_;
  uint64_t version_ = 0;
  uint64_t groups_ = 0;
  static constexpr int kLanes = QuerySlot::kLanes;
  // plain mode: one lane set on the live graph
  std::unique_ptr<Traversal> trav_[kLanes];
  DevBuf<uint32_t> stage_s_, stage_d_, stage_o_;
  DevBuf<uint32_t> res32_[kLanes];
  DevBuf<int32_t> par_;
  DevBuf<double> res64_;
  // batched mode: query groups run on snapshots while the live graph moves on
  std::vector<std::unique_ptr<QuerySlot>> slots_;
  size_t next_slot_ = 0;
  double t0_ = 0;

  // progress on stderr every tenth of the stream (TG_SERVE_PROGRESS=1)
  bool progress_ = false;
  std::atomic<uint64_t> completed_{0};
  void note_done(uint64_t k);

  double now() const;
  uint64_t footprint(const Request &r) const;
  void exec_update_group(const std::vector<uint64_t> &ids);
  void exec_query_plain(Request &r);
  void exec_query_group(QuerySlot &slot, std::vector<uint64_t> ids, uint64_t ver, uint64_t grp);
  QuerySlot &acquire_slot();
  QuerySlot *free_slot();
  void launch_group(QuerySlot &sl, std::vector<uint64_t> ids);
  void reject(Request &r);
  void run_admission(const std::vector<uint64_t> &order);
  void drain();
};

}  // namespace tg
