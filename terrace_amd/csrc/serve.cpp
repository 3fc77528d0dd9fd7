// Request loop: plain (one at a time) and batched (continuous batching) modes.
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>

namespace tg {

namespace {

double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  double pos = p * (v.size() - 1);
  size_t lo = static_cast<size_t>(std::floor(pos));
  size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - lo);
}

}  // namespace

Server::Server(Graph &g, int mode, uint64_t mem_budget, bool keep_results)
    : g_(g), mode_(mode), batched_(mode != kServePlain), budget_(mem_budget), keep_(keep_results) {
  if (mode < kServePlain || mode > kServeAdmission) throw Error(-1, "unknown serving mode");
  const char *pv = std::getenv("TG_SERVE_PROGRESS");
  progress_ = pv && pv[0] == '1';
  for (int i = 0; i < kLanes; i++) trav_[i] = std::make_unique<Traversal>(g_);
  if (batched_) {
    // snapshots in flight: bounded by the budget (each holds a full copy).
    // Admission mode keeps two, so the pending pool grows into large
    // shared groups while one group runs.
    uint64_t per = std::max<uint64_t>(g_.memory_bytes(), 1);
    uint64_t cap = mode_ == kServeAdmission ? 2 : 6;
    size_t k = static_cast<size_t>(std::min<uint64_t>(cap, std::max<uint64_t>(1, budget_ / (2 * per))));
    for (size_t j = 0; j < k; j++) {
      auto sl = std::make_unique<QuerySlot>();
      sl->snap = std::make_unique<Graph>(g_.num_vertices());
      for (int i = 0; i < kLanes; i++) {
        TG_HIP(hipStreamCreateWithFlags(&sl->streams[i], hipStreamNonBlocking));
        sl->trav[i] = std::make_unique<Traversal>(*sl->snap, sl->streams[i]);
      }
      slots_.push_back(std::move(sl));
    }
  }
  reserve();
}

// Lanes allocate their working memory (frontier queues and bitmaps, the
// sliced adjacency copy of PageRank, level and result arrays, the
// snapshot's own arrays) on first use and grow it by a quarter when a
// later call needs more.  On this device hipMalloc of a few GiB can take
// hundreds of milliseconds while earlier frees are reclaimed, and hipFree
// waits for the whole device, so a lane allocating mid-stream stalls every
// other lane.  Each lane therefore runs one query of its kind here, on the
// graph as it is: BFS from up to 64 sources among the vertices with the
// longest lists, CC, one PageRank iteration and one local triangle count.
// What those queries leave cached (tier lists, tree tasks, PageRank
// slices) is dropped afterwards, so the first requests still pay for
// their own work.
void Server::reserve() {
  const uint32_t n = g_.num_vertices();
  if (n == 0) return;
  auto probes = [n](Graph &g, hipStream_t s, std::vector<uint32_t> &src, uint32_t &ltc_v) {
    const Graph::TierLists &L = g.tier_lists();
    src.clear();
    uint32_t nt = std::min<uint32_t>(L.ntree, 64);
    uint32_t np = std::min<uint32_t>(L.npma, 64 - nt);
    src.resize(nt + np);
    if (nt) TG_HIP(hipMemcpyAsync(src.data(), L.tree, nt * 4ull, hipMemcpyDeviceToHost, s));
    if (np) TG_HIP(hipMemcpyAsync(src.data() + nt, L.pma, np * 4ull, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    if (src.empty()) src.push_back(0);
    // the triangle count probes a PMA vertex (moderate degree): a hub's
    // count can take far longer than the memory it reserves is worth
    ltc_v = np ? src[nt] : src[0];
  };
  std::vector<uint32_t> src;
  uint32_t lv = 0;
  uint64_t c = 0;
  if (!batched_) {
    Traversal &T = *trav_[0];
    hipStream_t s = T.stream();
    probes(g_, s, src, lv);
    par_.ensure(n, s);
    res32_[0].ensure(n, s);
    T.bfs(src[0], par_.get(), res32_[0].get());
    (void)T.digest(res32_[0].get(), nullptr, n);
    res32_[1].ensure(n, s);
    T.cc(res32_[1].get());
    res64_.ensure(n, s);
    T.pagerank(1, res64_.get());
    T.ltc(&lv, 1, &c);
    TG_HIP(hipStreamSynchronize(s));
    g_.touch();
    return;
  }
  for (auto &p : slots_) {
    QuerySlot &sl = *p;
    sl.snap->copy_from(g_, true);
    probes(*sl.snap, sl.streams[0], src, lv);
    const uint32_t k = static_cast<uint32_t>(src.size());
    {
      Traversal &T = *sl.trav[0];
      // a group runs its BFS requests 64 sources per call
      sl.res32[0].ensure(64ull * n, T.stream());
      T.bfs_multi(src.data(), k, sl.res32[0].get());
      (void)T.digest(sl.res32[0].get(), nullptr, n);
    }
    {
      Traversal &T = *sl.trav[1];
      sl.res32[1].ensure(n, T.stream());
      T.cc(sl.res32[1].get());
      (void)T.digest(sl.res32[1].get(), nullptr, n);
    }
    {
      Traversal &T = *sl.trav[2];
      sl.res64.ensure(n, T.stream());
      T.pagerank(1, sl.res64.get());
      (void)T.digest(nullptr, reinterpret_cast<const uint64_t *>(sl.res64.get()), n);
    }
    sl.trav[3]->ltc(&lv, 1, &c);
    for (int i = 0; i < kLanes; i++) TG_HIP(hipStreamSynchronize(sl.streams[i]));
    sl.snap->touch();
  }
  g_.touch();
}

Server::~Server() {
  try {
    finish();  // a loop still running on its own thread ends first
  } catch (...) {
  }
  try {
    drain();
  } catch (...) {
  }
  for (auto &sl : slots_) {
    for (int i = 0; i < kLanes; i++) {
      sl->trav[i].reset();
      if (sl->streams[i]) {
        (void)hipStreamSynchronize(sl->streams[i]);
        (void)hipStreamDestroy(sl->streams[i]);
      }
    }
  }
}

// A slot whose group has finished (joined), or null when all are busy.
QuerySlot *Server::free_slot() {
  for (auto &p : slots_) {
    QuerySlot &sl = *p;
    if (sl.running.load(std::memory_order_acquire)) continue;
    if (sl.worker.joinable()) sl.worker.join();
    if (!sl.error.empty()) throw Error(-3, "query group: " + sl.error);
    return &sl;
  }
  return nullptr;
}

// Snapshot the live graph into the slot and run the group on it.
void Server::launch_group(QuerySlot &sl, std::vector<uint64_t> ids) {
  sl.snap->copy_from(g_, true);  // version_ as of now, on the live graph's stream
  uint64_t ver = version_, grp = groups_++;
  sl.error.clear();
  sl.running.store(true, std::memory_order_release);
  sl.worker = std::thread([this, &sl, ids = std::move(ids), ver, grp] {
    try {
      exec_query_group(sl, ids, ver, grp);
    } catch (const std::exception &e) {
      sl.error = e.what();
    }
    sl.running.store(false, std::memory_order_release);
  });
}

void Server::reject(Request &r) {
  r.start = r.done = now();
  r.version = version_;
  publish(r, -2);
  note_done(1);
}

QuerySlot &Server::acquire_slot() {
  QuerySlot &sl = *slots_[next_slot_];
  next_slot_ = (next_slot_ + 1) % slots_.size();
  if (sl.worker.joinable()) sl.worker.join();
  if (!sl.error.empty()) throw Error(-3, "query group: " + sl.error);
  return sl;
}

void Server::drain() {
  std::string err;
  for (auto &sl : slots_) {
    if (sl->worker.joinable()) sl->worker.join();
    if (!sl->error.empty()) err = sl->error;
  }
  if (!err.empty()) throw Error(-3, "query group: " + err);
}

double Server::now() const { return wall() - t0_; }

void Server::note_done(uint64_t k) {
  uint64_t before = completed_.fetch_add(k);
  if (!progress_) return;
  uint64_t total;
  {
    std::lock_guard<std::mutex> l(mu_);
    total = total_;
  }
  uint64_t step = std::max<uint64_t>(1, total / 10);
  if ((before + k) / step != before / step || before + k == total)
    std::fprintf(stderr, "[serve] mode %d: %llu / %llu requests done at %.2f s\n", mode_,
                 static_cast<unsigned long long>(before + k), static_cast<unsigned long long>(total), now());
}

uint64_t Server::submit(Request r) {
  uint64_t id;
  {
    std::lock_guard<std::mutex> l(mu_);
    if (running_) {
      if (closed_) throw Error(-1, "the server's input is closed while its loop runs");
      // a live request arrives now (or at its own offset, if later): the
      // loop has taken in only requests that arrived before this moment
      r.arrival = std::max(r.arrival, now());
    }
    r.status = 1;
    reqs_.push_back(std::move(r));
    id = reqs_.size() - 1;
    total_++;
    if (running_) {
      inbox_.push_back(id);
      input_seq_++;
    }
  }
  input_cv_.notify_all();
  return id;
}

Request &Server::req(uint64_t id) {
  std::lock_guard<std::mutex> l(mu_);
  return reqs_.at(id);
}

// Results are written before the status leaves pending, under mu_: a
// reader that sees the status (under mu_) sees the results.
void Server::publish(Request &r, int status) {
  {
    std::lock_guard<std::mutex> l(mu_);
    r.status = status;
  }
  done_cv_.notify_all();
}

int Server::status(uint64_t id) const {
  std::lock_guard<std::mutex> l(mu_);
  if (id >= reqs_.size()) throw Error(-1, "unknown request id");
  return reqs_[id].status;
}

const Request &Server::result(uint64_t id) const {
  std::lock_guard<std::mutex> l(mu_);
  if (id >= reqs_.size()) throw Error(-1, "unknown request id");
  return reqs_[id];
}

bool Server::busy() const {
  std::lock_guard<std::mutex> l(mu_);
  return running_ || loop_.joinable();
}

int Server::wait(uint64_t id, double timeout_s) {
  std::unique_lock<std::mutex> l(mu_);
  if (id >= reqs_.size()) throw Error(-1, "unknown request id");
  const Request &r = reqs_[id];
  done_cv_.wait_for(l, std::chrono::duration<double>(std::max(0.0, timeout_s)),
                    [&] { return r.status != 1 || !running_; });
  return r.status;
}

// Device bytes a request needs while it runs (scratch, outputs, growth).
uint64_t Server::footprint(const Request &r) const {
  const uint64_t n = g_.num_vertices();
  switch (r.type) {
    case kReqUpdate: {
      uint64_t e = r.us.size();
      // keys, sort buffers, flags and kicked lists; PMA growth when the
      // array may double (16 B per slot of the current array)
      return 64 * e + (e > g_.pma_slots() / 8 ? 16 * g_.pma_slots() : 0);
    }
    case kReqBfs: return 4 * n + n / 2;   // one lane of a multi-source run
    case kReqCc: return 12 * n;
    // ranks and scratch (24 B per vertex), and the sliced adjacency copy
    // PageRank builds per graph version: 4 B per edge, 8 row-offset words
    // and 8 partial sums per vertex, and the degrees
    case kReqPr: return 24 * n + 4 * g_.edge_bound() + 100 * n;
    case kReqLtc: return 1 << 16;
    default: return 0;
  }
}

void Server::exec_update_group(const std::vector<uint64_t> &ids) {
  hipStream_t s = g_.stream();
  uint64_t total = 0;
  for (uint64_t id : ids) total += req(id).us.size();
  double st = now();
  if (!batched_) {
    // plain: the request's own inserts, then its deletes
    Request &r = req(ids[0]);
    std::vector<uint32_t> is, id, ds, dd;
    for (size_t i = 0; i < r.us.size(); i++) {
      if (r.uop[i] == 0) {
        is.push_back(r.us[i]);
        id.push_back(r.ud[i]);
      } else {
        ds.push_back(r.us[i]);
        dd.push_back(r.ud[i]);
      }
    }
    stage_s_.ensure(std::max<uint64_t>(total, 1), s);
    stage_d_.ensure(std::max<uint64_t>(total, 1), s);
    if (!is.empty()) {
      TG_HIP(hipMemcpyAsync(stage_s_.get(), is.data(), is.size() * 4, hipMemcpyHostToDevice, s));
      TG_HIP(hipMemcpyAsync(stage_d_.get(), id.data(), id.size() * 4, hipMemcpyHostToDevice, s));
      g_.insert_batch(stage_s_.get(), stage_d_.get(), is.size());
    }
    if (!ds.empty()) {
      TG_HIP(hipMemcpyAsync(stage_s_.get(), ds.data(), ds.size() * 4, hipMemcpyHostToDevice, s));
      TG_HIP(hipMemcpyAsync(stage_d_.get(), dd.data(), dd.size() * 4, hipMemcpyHostToDevice, s));
      g_.delete_batch(stage_s_.get(), stage_d_.get(), ds.size());
    }
  } else {
    std::vector<uint32_t> hs, hd, ho;
    hs.reserve(total);
    hd.reserve(total);
    ho.reserve(total);
    for (uint64_t rid : ids) {
      const Request &r = req(rid);
      hs.insert(hs.end(), r.us.begin(), r.us.end());
      hd.insert(hd.end(), r.ud.begin(), r.ud.end());
      ho.insert(ho.end(), r.uop.begin(), r.uop.end());
    }
    stage_s_.ensure(std::max<uint64_t>(total, 1), s);
    stage_d_.ensure(std::max<uint64_t>(total, 1), s);
    stage_o_.ensure(std::max<uint64_t>(total, 1), s);
    if (total) {
      TG_HIP(hipMemcpyAsync(stage_s_.get(), hs.data(), total * 4, hipMemcpyHostToDevice, s));
      TG_HIP(hipMemcpyAsync(stage_d_.get(), hd.data(), total * 4, hipMemcpyHostToDevice, s));
      TG_HIP(hipMemcpyAsync(stage_o_.get(), ho.data(), total * 4, hipMemcpyHostToDevice, s));
      g_.apply_ops(stage_s_.get(), stage_d_.get(), stage_o_.get(), total);
    }
  }
  TG_HIP(hipStreamSynchronize(s));
  // versions count update requests: the k-th batch of a coalesced run is
  // version base+k+1, and no query runs between them
  double t = now();
  for (uint64_t id : ids) {
    Request &r = req(id);
    r.start = st;
    r.done = t;
    r.version = ++version_;
    r.group = groups_;
    publish(r, 0);
  }
  groups_++;
  note_done(ids.size());
}

void Server::exec_query_plain(Request &r) {
  const uint32_t n = g_.num_vertices();
  Traversal &T = *trav_[0];
  hipStream_t s = T.stream();
  r.start = now();
  switch (r.type) {
    case kReqBfs: {
      par_.ensure(n, s);
      res32_[0].ensure(n, s);
      T.bfs(r.src, par_.get(), res32_[0].get());
      r.digest = T.digest(res32_[0].get(), nullptr, n);
      if (keep_) {
        r.r32.resize(n);
        TG_HIP(hipMemcpy(r.r32.data(), res32_[0].get(), n * 4ull, hipMemcpyDeviceToHost));
      }
      break;
    }
    case kReqCc: {
      res32_[1].ensure(n, s);
      T.cc(res32_[1].get());
      r.digest = T.digest(res32_[1].get(), nullptr, n);
      if (keep_) {
        r.r32.resize(n);
        TG_HIP(hipMemcpy(r.r32.data(), res32_[1].get(), n * 4ull, hipMemcpyDeviceToHost));
      }
      break;
    }
    case kReqPr: {
      res64_.ensure(n, s);
      T.pagerank(r.iters, res64_.get());
      r.digest = T.digest(nullptr, reinterpret_cast<const uint64_t *>(res64_.get()), n);
      std::vector<double> h(n);
      TG_HIP(hipMemcpy(h.data(), res64_.get(), n * 8ull, hipMemcpyDeviceToHost));
      double sum = 0;
      for (double x : h) sum += x;
      r.fvalue = sum;
      if (keep_) r.r64 = std::move(h);
      break;
    }
    case kReqLtc: {
      uint64_t c = 0;
      T.ltc(&r.src, 1, &c);
      r.value = c;
      break;
    }
    default:
      throw Error(-1, "unknown request type");
  }
  r.done = now();
  r.version = version_;
  r.group = groups_++;
  publish(r, 0);
  note_done(1);
}

void Server::exec_query_group(QuerySlot &slot, std::vector<uint64_t> ids, uint64_t ver, uint64_t grp) {
  const uint32_t n = g_.num_vertices();
  std::vector<uint64_t> bfs, cc, pr, ltc;
  for (uint64_t id : ids) {
    switch (req(id).type) {
      case kReqBfs: bfs.push_back(id); break;
      case kReqCc: cc.push_back(id); break;
      case kReqPr: pr.push_back(id); break;
      case kReqLtc: ltc.push_back(id); break;
      default: throw Error(-1, "unknown request type");
    }
  }
  (void)slot.snap->tier_lists();  // shared, read-only while the group runs
  double st = now();
  auto finish = [&](uint64_t id) {
    Request &r = req(id);
    r.start = st;
    r.done = now();
    r.version = ver;
    r.group = grp;
    publish(r, 0);
  };
  std::vector<std::thread> th;
  std::vector<std::string> errors(kLanes);
  if (!bfs.empty())
    th.emplace_back([&] {
      try {
        Traversal &T = *slot.trav[0];
        hipStream_t s = T.stream();
        for (size_t c0 = 0; c0 < bfs.size(); c0 += 64) {
          uint32_t k = static_cast<uint32_t>(std::min<size_t>(64, bfs.size() - c0));
          std::vector<uint32_t> srcs(k);
          for (uint32_t j = 0; j < k; j++) srcs[j] = req(bfs[c0 + j]).src;
          slot.res32[0].ensure(static_cast<uint64_t>(k) * n, s);
          T.bfs_multi(srcs.data(), k, slot.res32[0].get());
          for (uint32_t j = 0; j < k; j++) {
            Request &r = req(bfs[c0 + j]);
            const uint32_t *lv = slot.res32[0].get() + static_cast<uint64_t>(j) * n;
            r.digest = T.digest(lv, nullptr, n);
            if (keep_) {
              r.r32.resize(n);
              TG_HIP(hipMemcpyAsync(r.r32.data(), lv, n * 4ull, hipMemcpyDeviceToHost, s));
              TG_HIP(hipStreamSynchronize(s));
            }
            finish(bfs[c0 + j]);
          }
        }
      } catch (const std::exception &e) {
        errors[0] = e.what();
      }
    });
  if (!cc.empty())
    th.emplace_back([&] {
      try {
        Traversal &T = *slot.trav[1];
        hipStream_t s = T.stream();
        slot.res32[1].ensure(n, s);
        T.cc(slot.res32[1].get());
        uint64_t d = T.digest(slot.res32[1].get(), nullptr, n);
        std::vector<uint32_t> h;
        if (keep_) {
          h.resize(n);
          TG_HIP(hipMemcpyAsync(h.data(), slot.res32[1].get(), n * 4ull, hipMemcpyDeviceToHost, s));
          TG_HIP(hipStreamSynchronize(s));
        }
        for (uint64_t id : cc) {
          Request &r = req(id);
          r.digest = d;
          if (keep_) r.r32 = h;
          finish(id);
        }
      } catch (const std::exception &e) {
        errors[1] = e.what();
      }
    });
  if (!pr.empty())
    th.emplace_back([&] {
      try {
        Traversal &T = *slot.trav[2];
        hipStream_t s = T.stream();
        std::map<uint32_t, std::vector<uint64_t>> by_iters;
        for (uint64_t id : pr) by_iters[req(id).iters].push_back(id);
        slot.res64.ensure(n, s);
        for (auto &kv : by_iters) {
          T.pagerank(kv.first, slot.res64.get());
          uint64_t d = T.digest(nullptr, reinterpret_cast<const uint64_t *>(slot.res64.get()), n);
          std::vector<double> h(n);
          TG_HIP(hipMemcpyAsync(h.data(), slot.res64.get(), n * 8ull, hipMemcpyDeviceToHost, s));
          TG_HIP(hipStreamSynchronize(s));
          double sum = 0;
          for (double x : h) sum += x;
          for (uint64_t id : kv.second) {
            Request &r = req(id);
            r.digest = d;
            r.fvalue = sum;
            if (keep_) r.r64 = h;
            finish(id);
          }
        }
      } catch (const std::exception &e) {
        errors[2] = e.what();
      }
    });
  if (!ltc.empty())
    th.emplace_back([&] {
      try {
        Traversal &T = *slot.trav[3];
        std::vector<uint32_t> vs(ltc.size());
        for (size_t i = 0; i < ltc.size(); i++) vs[i] = req(ltc[i]).src;
        std::vector<uint64_t> out(ltc.size());
        T.ltc(vs.data(), static_cast<uint32_t>(vs.size()), out.data());
        for (size_t i = 0; i < ltc.size(); i++) {
          req(ltc[i]).value = out[i];
          finish(ltc[i]);
        }
      } catch (const std::exception &e) {
        errors[3] = e.what();
      }
    });
  for (auto &t : th) t.join();
  for (auto &e : errors)
    if (!e.empty()) throw Error(-3, "query group: " + e);
  note_done(ids.size());
}

void Server::begin(bool live) {
  std::lock_guard<std::mutex> l(mu_);
  if (running_) throw Error(-1, "the server's loop is already running");
  running_ = true;
  closed_ = !live;
  loop_code_ = 0;
  loop_msg_.clear();
  inbox_.clear();
  for (uint64_t i = 0; i < reqs_.size(); i++)
    if (reqs_[i].status == 1) inbox_.push_back(i);
  input_seq_++;
  completed_ = 0;
  t0_ = wall();
}

void Server::end() {
  {
    std::lock_guard<std::mutex> l(mu_);
    running_ = false;
    closed_ = true;
  }
  done_cv_.notify_all();
}

void Server::run() {
  if (loop_.joinable()) throw Error(-1, "the server's loop is running on its own thread: finish() it first");
  begin(false);
  try {
    loop_body();
  } catch (...) {
    end();
    throw;
  }
  end();
}

void Server::start(std::mutex *hold) {
  if (loop_.joinable()) throw Error(-1, "the server's loop is already running");
  begin(true);
  loop_ = std::thread([this, hold] {
    std::unique_lock<std::mutex> graph_lock;
    if (hold) graph_lock = std::unique_lock<std::mutex>(*hold);
    int code = 0;
    std::string msg;
    try {
      loop_body();
    } catch (const Error &e) {
      code = e.code;
      msg = e.what();
    } catch (const std::exception &e) {
      code = -4;
      msg = e.what();
    }
    {
      std::lock_guard<std::mutex> l(mu_);
      loop_code_ = code;
      loop_msg_ = msg;
    }
    end();
  });
}

void Server::finish() {
  {
    std::lock_guard<std::mutex> l(mu_);
    closed_ = true;
    input_seq_++;
  }
  input_cv_.notify_all();
  if (loop_.joinable()) loop_.join();
  int code;
  std::string msg;
  {
    std::lock_guard<std::mutex> l(mu_);
    code = loop_code_;
    msg = loop_msg_;
    loop_code_ = 0;
  }
  if (code) throw Error(code, "server loop: " + msg);
}

bool Server::take_input(std::vector<uint64_t> &order, size_t next, double &t, uint64_t &seen) {
  std::lock_guard<std::mutex> l(mu_);
  seen = input_seq_;
  if (!inbox_.empty()) {
    for (uint64_t id : inbox_) {
      reqs_[id].footprint = footprint(reqs_[id]);
      order.push_back(id);
    }
    inbox_.clear();
    std::stable_sort(order.begin() + static_cast<std::ptrdiff_t>(next), order.end(),
                     [&](uint64_t a, uint64_t b) { return reqs_[a].arrival < reqs_[b].arrival; });
  }
  // taken under mu_: a request submitted after this arrives at t or later
  t = now();
  return closed_;
}

void Server::wait_input(uint64_t seen, double seconds) {
  std::unique_lock<std::mutex> l(mu_);
  input_cv_.wait_for(l, std::chrono::duration<double>(seconds), [&] { return input_seq_ != seen; });
}

void Server::loop_body() {
  (void)g_.tier_lists();
  if (mode_ == kServeAdmission) run_admission();
  else run_strict();
}

void Server::run_strict() {
  std::vector<uint64_t> order;  // requests taken in, in arrival order
  size_t next = 0;              // first one not yet executed
  uint64_t seen = 0;
  while (true) {
    double t = 0;
    const bool closed = take_input(order, next, t, seen);
    if (next >= order.size()) {
      if (closed) break;
      wait_input(seen, 1.0);
      continue;
    }
    Request &head = req(order[next]);
    if (head.arrival > t) {
      double wait = head.arrival - t;
      if (wait > 200e-6) wait_input(seen, wait - 100e-6);
      continue;
    }
    // a request larger than the whole budget can never be admitted
    if (head.footprint > budget_) {
      reject(head);
      next++;
      continue;
    }
    if (!batched_) {
      if (head.type == kReqUpdate) exec_update_group({order[next]});
      else exec_query_plain(head);
      next++;
      continue;
    }
    // maximal arrival-ordered run of the head's kind that has arrived and fits
    bool upd = head.type == kReqUpdate;
    std::vector<uint64_t> ids;
    uint64_t used = 0;
    size_t j = next;
    while (j < order.size()) {
      Request &r = req(order[j]);
      if (r.arrival > t || (r.type == kReqUpdate) != upd) break;
      if (r.footprint > budget_) break;  // rejected on its own turn
      if (!ids.empty() && used + r.footprint > budget_) break;
      used += r.footprint;
      ids.push_back(order[j]);
      j++;
    }
    if (upd) {
      exec_update_group(ids);
    } else {
      launch_group(acquire_slot(), ids);
    }
    next = j;
  }
  drain();
}

void Server::run_admission() {
  std::vector<uint64_t> order;   // requests taken in, in arrival order
  size_t next = 0;               // first one not yet taken from order
  std::vector<uint64_t> pool;    // arrived queries waiting for a slot, in arrival order
  uint64_t seen = 0;
  while (true) {
    double t = 0;
    const bool closed = take_input(order, next, t, seen);
    if (closed && next >= order.size() && pool.empty()) break;
    // take in every arrival: updates are applied now, in arrival order
    std::vector<uint64_t> upd;
    uint64_t used_u = 0;
    while (next < order.size() && req(order[next]).arrival <= t) {
      uint64_t id = order[next++];
      Request &r = req(id);
      if (r.footprint > budget_) {
        reject(r);
        continue;
      }
      if (r.type != kReqUpdate) {
        pool.push_back(id);
        continue;
      }
      if (!upd.empty() && used_u + r.footprint > budget_) {
        exec_update_group(upd);
        upd.clear();
        used_u = 0;
      }
      upd.push_back(id);
      used_u += r.footprint;
    }
    if (!upd.empty()) exec_update_group(upd);
    // admit the pool (oldest first, within the budget) when a slot is free
    if (!pool.empty()) {
      if (QuerySlot *sl = free_slot()) {
        std::vector<uint64_t> ids;
        uint64_t used = 0;
        size_t k = 0;
        for (; k < pool.size(); k++) {
          const Request &r = req(pool[k]);
          if (!ids.empty() && used + r.footprint > budget_) break;
          used += r.footprint;
          ids.push_back(pool[k]);
        }
        pool.erase(pool.begin(), pool.begin() + static_cast<std::ptrdiff_t>(k));
        launch_group(*sl, std::move(ids));
        continue;
      }
    }
    // idle: wait for the next arrival or new input, or poll for a slot
    double wait = next < order.size() ? req(order[next]).arrival - now() : 1.0;
    if (!pool.empty()) wait = std::min(wait, 50e-6);
    if (wait > 20e-6) wait_input(seen, std::min(wait, 1e-3));
  }
  drain();
}

std::vector<double> Server::summary() const {
  std::lock_guard<std::mutex> l(mu_);
  std::vector<double> all, up, q;
  double first = 1e300, last = 0;
  double rejected = 0;
  for (const Request &r : reqs_) {
    if (r.status < 0) rejected++;
    if (r.status != 0) continue;
    double lat = r.done - r.arrival;
    all.push_back(lat);
    (r.type == kReqUpdate ? up : q).push_back(lat);
    first = std::min(first, r.arrival);
    last = std::max(last, r.done);
  }
  double span = all.empty() ? 0 : last - first;
  return {span, span > 0 ? all.size() / span : 0, pct(all, 0.5), pct(all, 0.99), pct(up, 0.5), pct(up, 0.99),
          pct(q, 0.5), pct(q, 0.99), static_cast<double>(groups_), rejected};
}

}  // namespace tg
