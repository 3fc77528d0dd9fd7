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
//
// run() replays the requests submitted before it and returns when they are
// done.  start() runs the same loop on a thread of its own and keeps the
// input open: a request submitted while the loop runs arrives at the time
// of its submission (or at its own arrival offset, if later), and is
// ordered, batched and versioned like the rest; finish() closes the input,
// waits for every request and joins the loop.
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
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
  // Thread-safe; callable before run()/start() and while the loop runs.
  uint64_t submit(Request r);
  void run();
  // The loop on its own thread, holding *hold (if not null) while it runs.
  void start(std::mutex *hold);
  // Close the input, wait for every request, join the loop thread; throws
  // the loop's error, if it failed.
  void finish();
  // Status of request id (1 pending, 0 done, < 0 error) after waiting up to
  // timeout_s for it to leave pending (or for the loop to end).
  int wait(uint64_t id, double timeout_s);
  // The loop is running (run() or start() until finish()).  run(), start()
  // and finish() are called from one controlling thread.
  bool busy() const;
  int status(uint64_t id) const;
  // A request's fields; stable once its status is no longer pending.
  const Request &result(uint64_t id) const;
  // [0] makespan s, [1] requests/s, [2] p50 s, [3] p99 s, [4] p50 update,
  // [5] p99 update, [6] p50 query, [7] p99 query, [8] groups, [9] rejected
  // (completed requests only)
  std::vector<double> summary() const;

 private:
  Graph &g_;
  int mode_;
  bool batched_;
  uint64_t budget_;
  bool keep_;
  // Requests by id.  A deque keeps references valid while submit() appends;
  // mu_ guards the container, the input list and every status change.
  std::deque<Request> reqs_;
  mutable std::mutex mu_;
  std::condition_variable input_cv_;  // new input or closed input
  mutable std::condition_variable done_cv_;  // a request completed, or the loop ended
  std::vector<uint64_t> inbox_;       // submitted, not yet taken in by the loop
  uint64_t input_seq_ = 0;            // bumped by every submit and by finish()
  bool running_ = false, closed_ = true;
  std::thread loop_;
  int loop_code_ = 0;
  std::string loop_msg_;
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
  uint64_t total_ = 0;  // requests submitted (under mu_)
  void note_done(uint64_t k);

  double now() const;
  Request &req(uint64_t id);
  void publish(Request &r, int status);
  void begin(bool live);
  void end();
  void loop_body();
  // Takes the submitted requests into order (the unexecuted tail, from
  // next, kept in arrival order); t = the server clock at that moment.
  // Returns true when the input is closed.
  bool take_input(std::vector<uint64_t> &order, size_t next, double &t, uint64_t &seen);
  void wait_input(uint64_t seen, double seconds);
  uint64_t footprint(const Request &r) const;
  void exec_update_group(const std::vector<uint64_t> &ids);
  void exec_query_plain(Request &r);
  void exec_query_group(QuerySlot &slot, std::vector<uint64_t> ids, uint64_t ver, uint64_t grp);
  QuerySlot &acquire_slot();
  QuerySlot *free_slot();
  void launch_group(QuerySlot &sl, std::vector<uint64_t> ids);
  void reject(Request &r);
  // Start-up: every lane runs one query of its kind, so each allocates
  // its working memory before the first request (see serve.cpp)
  void reserve();
  void run_strict();
  void run_admission();
  void drain();
};

}  // namespace tg
