// extern "C" boundary of libterrace_hip.so (declared in include/terrace_hip.h).
// Every entry point catches all C++ exceptions and returns a status code.
#include "../../include/terrace_hip.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "gen.hpp"
#include "io.hpp"
#include "serve.hpp"
#include "traverse.hpp"

struct tg_graph {
  tg::Graph g;
  tg::Traversal t;
  // host-call staging buffers
  tg::DevBuf<uint32_t> s_src, s_dst, s_u32;
  tg::DevBuf<uint8_t> s_u8;
  tg::DevBuf<uint64_t> s_u64;
  tg::DevBuf<int32_t> s_i32;
  tg::DevBuf<double> s_f64;
  int device = 0;  // the device current when the graph was created
  // profiled update steps: one span of stream events from the first step
  // after a reset to the end of the latest, over `steps` calls
  hipEvent_t span_a = nullptr, span_b = nullptr;
  uint64_t span_steps = 0;
  explicit tg_graph(uint32_t n) : g(n), t(g) { TG_HIP(hipGetDevice(&device)); }
  ~tg_graph() {
    if (span_a) (void)hipEventDestroy(span_a);
    if (span_b) (void)hipEventDestroy(span_b);
  }
};

struct tg_frontier {
  tg_graph *owner;
  tg::Frontier f;
};

namespace {

thread_local std::string g_err;

template <class F>
int guard(F &&f) {
  try {
    f();
    return TG_OK;
  } catch (const tg::Error &e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_err = "host allocation failed";
    return TG_ENOMEM;
  } catch (const std::exception &e) {
    g_err = e.what();
    return TG_EINTERNAL;
  } catch (...) {
    g_err = "unknown error";
    return TG_EINTERNAL;
  }
}

#define NEED(cond, msg) \
  if (!(cond)) throw tg::Error(TG_EINVAL, msg)

void h2d(void *d, const void *h, uint64_t bytes, hipStream_t s) {
  if (bytes) TG_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
}
void d2h(void *h, const void *d, uint64_t bytes, hipStream_t s) {
  if (bytes) TG_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
}

void stage_pair(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count) {
  hipStream_t s = G->g.stream();
  G->s_src.ensure(count + 1, s);
  G->s_dst.ensure(count + 1, s);
  h2d(G->s_src.get(), src, count * 4, s);
  h2d(G->s_dst.get(), dst, count * 4, s);
}

}  // namespace

extern "C" {

const char *tg_last_error(void) { return g_err.c_str(); }

int tg_create(uint32_t num_vertices, tg_graph **out) {
  return guard([&] {
    NEED(out, "null out");
    *out = nullptr;
    NEED(num_vertices > 0, "num_vertices must be > 0");
    *out = new tg_graph(num_vertices);
  });
}

int tg_create_weighted(uint32_t num_vertices, tg_graph **out) {
  return guard([&] {
    NEED(out, "null out");
    *out = nullptr;
    NEED(num_vertices > 0, "num_vertices must be > 0");
    std::unique_ptr<tg_graph> G(new tg_graph(num_vertices));
    G->g.enable_weights();
    *out = G.release();
  });
}

int tg_is_weighted(tg_graph *G, int *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    *out = G->g.weighted() ? 1 : 0;
  });
}

int tg_insert_batch_weighted(tg_graph *G, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                             uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst && w)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (!count) return;
    hipStream_t st = G->g.stream();
    stage_pair(G, src, dst, count);
    G->s_u32.ensure(count + 1, st);
    h2d(G->s_u32.get(), w, count * 4, st);
    G->g.insert_batch_weighted(G->s_src.get(), G->s_dst.get(), G->s_u32.get(), count);
  });
}

int tg_insert_batch_weighted_dev(tg_graph *G, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                                 uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst && w)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->g.insert_batch_weighted(src, dst, w, count);
  });
}

int tg_edge_weights(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count, uint32_t *out) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst && out)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(G->g.weighted(), "graph is not weighted");
    if (!count) return;
    hipStream_t st = G->g.stream();
    stage_pair(G, src, dst, count);
    G->s_u32.ensure(count + 1, st);
    G->g.edge_weights(G->s_src.get(), G->s_dst.get(), count, G->s_u32.get());
    d2h(out, G->s_u32.get(), count * 4, st);
  });
}

int tg_destroy(tg_graph *g) {
  return guard([&] { delete g; });
}

int tg_insert_batch(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (!count) return;
    stage_pair(G, src, dst, count);
    G->g.insert_batch(G->s_src.get(), G->s_dst.get(), count);
  });
}

int tg_delete_batch(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (!count) return;
    stage_pair(G, src, dst, count);
    G->g.delete_batch(G->s_src.get(), G->s_dst.get(), count);
  });
}

int tg_insert_batch_dev(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->g.insert_batch(src, dst, count);
  });
}

int tg_delete_batch_dev(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->g.delete_batch(src, dst, count);
  });
}

int tg_update_batch_dev(tg_graph *G, const uint32_t *ins_src, const uint32_t *ins_dst, uint64_t n_ins,
                        const uint32_t *del_src, const uint32_t *del_dst, uint64_t n_del) {
  return guard([&] {
    NEED(G && (n_ins == 0 || (ins_src && ins_dst)) && (n_del == 0 || (del_src && del_dst)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (!G->t.profiling) {
      G->g.update_batch(ins_src, ins_dst, n_ins, del_src, del_dst, n_del);
      return;
    }
    // profiled: consecutive steps on the graph's stream, as "update_step":
    // an event before the first step after a reset; tg_profile_get records
    // the closing one on the stream (nothing is recorded between steps)
    if (!G->span_a) {
      TG_HIP(hipEventCreate(&G->span_a));
      TG_HIP(hipEventRecord(G->span_a, G->g.stream()));
    }
    G->g.update_batch(ins_src, ins_dst, n_ins, del_src, del_dst, n_del);
    G->span_steps++;
  });
}

int tg_update_batch(tg_graph *G, const uint32_t *ins_src, const uint32_t *ins_dst, uint64_t n_ins,
                    const uint32_t *del_src, const uint32_t *del_dst, uint64_t n_del) {
  return guard([&] {
    NEED(G && (n_ins == 0 || (ins_src && ins_dst)) && (n_del == 0 || (del_src && del_dst)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    const uint64_t count = n_ins + n_del;
    if (!count) return;
    // staged as one array pair: the inserts, then the deletes
    hipStream_t st = G->g.stream();
    G->s_src.ensure(count + 1, st);
    G->s_dst.ensure(count + 1, st);
    h2d(G->s_src.get(), ins_src, n_ins * 4, st);
    h2d(G->s_dst.get(), ins_dst, n_ins * 4, st);
    h2d(G->s_src.get() + n_ins, del_src, n_del * 4, st);
    h2d(G->s_dst.get() + n_ins, del_dst, n_del * 4, st);
    G->g.update_batch(G->s_src.get(), G->s_dst.get(), n_ins, G->s_src.get() + n_ins, G->s_dst.get() + n_ins, n_del);
  });
}

static int single_edge(tg_graph *G, uint32_t s, uint32_t d, int *result, bool insert) {
  return guard([&] {
    NEED(G && result, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(s < G->g.num_vertices() && d < G->g.num_vertices(), "vertex id out of range");
    hipStream_t st = G->g.stream();
    stage_pair(G, &s, &d, 1);
    G->s_u8.ensure(2, st);
    G->g.is_edge_batch(G->s_src.get(), G->s_dst.get(), 1, G->s_u8.get());
    uint8_t present = 0;
    d2h(&present, G->s_u8.get(), 1, st);
    if (insert) {
      *result = present ? -1 : 0;
      if (!present) G->g.insert_batch(G->s_src.get(), G->s_dst.get(), 1);
    } else {
      *result = present ? 0 : -1;
      if (present) G->g.delete_batch(G->s_src.get(), G->s_dst.get(), 1);
    }
  });
}

int tg_add_edge(tg_graph *G, uint32_t s, uint32_t d, int *result) { return single_edge(G, s, d, result, true); }
int tg_remove_edge(tg_graph *G, uint32_t s, uint32_t d, int *result) { return single_edge(G, s, d, result, false); }

int tg_is_edge(tg_graph *G, uint32_t s, uint32_t d, uint32_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(s < G->g.num_vertices() && d < G->g.num_vertices(), "vertex id out of range");
    hipStream_t st = G->g.stream();
    stage_pair(G, &s, &d, 1);
    G->s_u8.ensure(2, st);
    G->g.is_edge_batch(G->s_src.get(), G->s_dst.get(), 1, G->s_u8.get());
    uint8_t r = 0;
    d2h(&r, G->s_u8.get(), 1, st);
    *out = r;
  });
}

int tg_is_edge_batch(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t count, uint8_t *out) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst && out)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (!count) return;
    hipStream_t st = G->g.stream();
    stage_pair(G, src, dst, count);
    G->s_u8.ensure(count + 1, st);
    G->g.is_edge_batch(G->s_src.get(), G->s_dst.get(), count, G->s_u8.get());
    d2h(out, G->s_u8.get(), count, st);
  });
}

int tg_count_common(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out) {
  return guard([&] {
    NEED(G && (q == 0 || (src && dst && out)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (!q) return;
    hipStream_t st = G->g.stream();
    stage_pair(G, src, dst, q);
    G->s_u32.ensure(q + 1, st);
    G->g.count_common(G->s_src.get(), G->s_dst.get(), q, G->s_u32.get());
    d2h(out, G->s_u32.get(), q * 4, st);
  });
}

int tg_count_common_dev(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out) {
  return guard([&] {
    NEED(G && (q == 0 || (src && dst && out)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->g.count_common(src, dst, q, out);
    TG_HIP(hipStreamSynchronize(G->g.stream()));
  });
}

// ---- no-allocation count_common
// The staging of tg_count_common_noalloc lives in a static device array
// (each GPU holds its own copy until the process ends): [0, 256) the id-check
// word, then src, dst and out, kNoAllocPairs uint32 each.  Host code never
// passes the array's name to a kernel (it is a host-side stand-in there):
// each device's address comes from hipGetSymbolAddress, taken once per
// device with that device current, and every launch and copy uses it.
namespace {
constexpr uint64_t kNoAllocPairs = 1ull << 20;
constexpr uint64_t kNoAllocBytes = 256 + 3 * 4 * kNoAllocPairs;  // 12 MiB + 256 B
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;  // one call at a time in this process: the calls share the array
unsigned char *g_noalloc_ptr[kMaxDevices] = {};

// Caller holds g_noalloc_mu.
unsigned char *noalloc_scratch(int device) {
  NEED(device >= 0 && device < kMaxDevices, "device id out of range");
  if (!g_noalloc_ptr[device]) {
    int cur = 0;
    TG_HIP(hipGetDevice(&cur));
    if (cur != device) TG_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) TG_HIP(hipSetDevice(cur));
    if (e != hipSuccess) throw tg::Error(TG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<unsigned char *>(p);
  }
  return g_noalloc_ptr[device];
}
}  // namespace

int tg_count_common_noalloc(tg_graph *G, const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out) {
  return guard([&] {
    NEED(G && (q == 0 || (src && dst && out)), "null argument");
    NEED(q <= kNoAllocPairs, "tg_count_common_noalloc: at most 2^20 pairs per call");
    if (!q) return;
    std::lock_guard<std::mutex> ls(g_noalloc_mu);
    std::lock_guard<std::mutex> l(G->g.mu);
    unsigned char *base = noalloc_scratch(G->device);
    uint64_t *err = reinterpret_cast<uint64_t *>(base);  // cleared by a kernel in count_common
    uint32_t *s = reinterpret_cast<uint32_t *>(base + 256);
    uint32_t *d = s + kNoAllocPairs, *o = d + kNoAllocPairs;
    hipStream_t st = G->g.stream();
    h2d(s, src, q * 4, st);
    h2d(d, dst, q * 4, st);
    G->g.count_common(s, d, q, o, err);
    d2h(out, o, q * 4, st);
  });
}

int tg_degree(tg_graph *G, uint32_t v, uint32_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(v < G->g.num_vertices(), "vertex id out of range");
    hipStream_t st = G->g.stream();
    uint32_t n = G->g.num_vertices();
    G->s_u32.ensure(n, st);
    G->g.degrees(G->s_u32.get());
    d2h(out, G->s_u32.get() + v, 4, st);
  });
}

int tg_degrees_all(tg_graph *G, uint32_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    hipStream_t st = G->g.stream();
    uint32_t n = G->g.num_vertices();
    G->s_u32.ensure(n, st);
    G->g.degrees(G->s_u32.get());
    d2h(out, G->s_u32.get(), static_cast<uint64_t>(n) * 4, st);
  });
}

int tg_num_edges(tg_graph *G, uint64_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    *out = G->g.num_edges();
  });
}

int tg_num_vertices(tg_graph *G, uint32_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    *out = G->g.num_vertices();
  });
}

int tg_memory_bytes(tg_graph *G, uint64_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    *out = G->g.memory_bytes();
  });
}

int tg_export_adjacency_dev(tg_graph *G, uint64_t *offsets, uint32_t *nbrs) {
  return guard([&] {
    NEED(G && offsets && nbrs, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->g.export_adjacency(offsets, nbrs);
  });
}

int tg_export_adjacency(tg_graph *G, uint64_t *offsets, uint32_t *nbrs) {
  return guard([&] {
    NEED(G && offsets, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    hipStream_t st = G->g.stream();
    uint32_t n = G->g.num_vertices();
    uint64_t m = G->g.num_edges();
    NEED(m == 0 || nbrs, "null neighbours");
    G->s_u64.ensure(static_cast<uint64_t>(n) + 1, st);
    G->s_u32.ensure(m + 1, st);
    G->g.export_adjacency(G->s_u64.get(), G->s_u32.get());
    d2h(offsets, G->s_u64.get(), (static_cast<uint64_t>(n) + 1) * 8, st);
    d2h(nbrs, G->s_u32.get(), m * 4, st);
  });
}

int tg_neighbors(tg_graph *G, uint32_t v, uint32_t *out, uint32_t cap, uint32_t *degree) {
  return guard([&] {
    NEED(G && degree, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(v < G->g.num_vertices(), "vertex id out of range");
    hipStream_t st = G->g.stream();
    uint32_t n = G->g.num_vertices();
    G->s_u32.ensure(n, st);
    G->g.degrees(G->s_u32.get());
    uint32_t deg = 0;
    d2h(&deg, G->s_u32.get() + v, 4, st);
    *degree = deg;
    if (!deg || !cap) return;
    NEED(out, "null out");
    G->s_dst.ensure(static_cast<uint64_t>(deg) + 64, st);
    G->g.neighbors_of(v, G->s_dst.get());
    d2h(out, G->s_dst.get(), static_cast<uint64_t>(std::min(deg, cap)) * 4, st);
  });
}

int tg_tiers(tg_graph *G, uint8_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    hipStream_t st = G->g.stream();
    uint32_t n = G->g.num_vertices();
    G->s_u8.ensure(n, st);
    G->g.tiers(G->s_u8.get());
    d2h(out, G->s_u8.get(), n, st);
  });
}

// ---------------------------------------------------------------- algorithms
int tg_bfs_dev(tg_graph *G, uint32_t src, int32_t *parents, uint32_t *levels, uint32_t *depth) {
  return guard([&] {
    NEED(G && parents, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(src < G->g.num_vertices(), "source out of range");
    uint32_t d = G->t.bfs(src, parents, levels);
    if (depth) *depth = d;
  });
}

int tg_bfs(tg_graph *G, uint32_t src, int32_t *parents, uint32_t *levels) {
  return guard([&] {
    NEED(G, "null graph");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    NEED(src < n, "source out of range");
    hipStream_t st = G->g.stream();
    G->s_i32.ensure(n, st);
    G->s_u32.ensure(n, st);
    G->t.bfs(src, G->s_i32.get(), G->s_u32.get());
    if (parents) d2h(parents, G->s_i32.get(), static_cast<uint64_t>(n) * 4, st);
    if (levels) d2h(levels, G->s_u32.get(), static_cast<uint64_t>(n) * 4, st);
  });
}

int tg_pagerank_dev(tg_graph *G, uint32_t iters, double *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->t.pagerank(iters, out);
  });
}

int tg_pagerank(tg_graph *G, uint32_t iters, double *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    hipStream_t st = G->g.stream();
    G->s_f64.ensure(n, st);
    G->t.pagerank(iters, G->s_f64.get());
    d2h(out, G->s_f64.get(), static_cast<uint64_t>(n) * 8, st);
  });
}

int tg_cc_dev(tg_graph *G, uint32_t *labels, uint32_t *rounds) {
  return guard([&] {
    NEED(G && labels, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t r = G->t.cc(labels);
    if (rounds) *rounds = r;
  });
}

int tg_cc(tg_graph *G, uint32_t *labels) {
  return guard([&] {
    NEED(G && labels, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    hipStream_t st = G->g.stream();
    G->s_u32.ensure(n, st);
    G->t.cc(G->s_u32.get());
    d2h(labels, G->s_u32.get(), static_cast<uint64_t>(n) * 4, st);
  });
}

int tg_bc_dev(tg_graph *G, uint32_t src, double *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(src < G->g.num_vertices(), "source out of range");
    G->t.bc(src, out);
  });
}

int tg_bc(tg_graph *G, uint32_t src, double *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    NEED(src < n, "source out of range");
    hipStream_t st = G->g.stream();
    G->s_f64.ensure(n, st);
    G->t.bc(src, G->s_f64.get());
    d2h(out, G->s_f64.get(), static_cast<uint64_t>(n) * 8, st);
  });
}

int tg_sssp_bf_dev(tg_graph *G, uint32_t src, int32_t *dist, uint32_t *rounds) {
  return guard([&] {
    NEED(G && dist, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(src < G->g.num_vertices(), "source out of range");
    uint32_t r = G->t.sssp_bf(src, dist);
    if (rounds) *rounds = r;
  });
}

int tg_sssp_bf(tg_graph *G, uint32_t src, int32_t *dist) {
  return guard([&] {
    NEED(G && dist, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    NEED(src < n, "source out of range");
    hipStream_t st = G->g.stream();
    G->s_i32.ensure(n, st);
    G->t.sssp_bf(src, G->s_i32.get());
    d2h(dist, G->s_i32.get(), static_cast<uint64_t>(n) * 4, st);
  });
}

int tg_tc(tg_graph *G, uint64_t *count) {
  return guard([&] {
    NEED(G && count, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    *count = G->t.tc();
  });
}

// ---------------------------------------------------------------- frontiers
int tg_frontier_create(tg_graph *G, tg_frontier **out) {
  return guard([&] {
    NEED(G && out, "null argument");
    auto *f = new tg_frontier;
    f->owner = G;
    f->f.n = G->g.num_vertices();
    f->f.count = 0;
    f->f.has_queue = true;
    f->f.queue.ensure(f->f.n, G->g.stream());
    *out = f;
  });
}

int tg_frontier_create_n(uint32_t n, tg_frontier **out) {
  return guard([&] {
    NEED(out && n, "null argument or n = 0");
    auto *f = new tg_frontier;
    f->owner = nullptr;  // bound to a graph by its first edge map
    f->f.n = n;
    f->f.count = 0;
    f->f.has_queue = true;
    f->f.queue.ensure(n, nullptr);
    *out = f;
  });
}

int tg_frontier_destroy(tg_frontier *f) {
  return guard([&] { delete f; });
}

namespace {
// The frontier's graph lock and stream; an unbound frontier uses the null
// stream with synchronous copies.
struct FrontierCtx {
  std::unique_lock<std::mutex> lk;
  hipStream_t st = nullptr;
  explicit FrontierCtx(tg_frontier *F) {
    if (F->owner) {
      lk = std::unique_lock<std::mutex>(F->owner->g.mu);
      st = F->owner->g.stream();
    }
  }
};

// Binds an unbound frontier to G (same vertex count); false when F belongs
// to another graph.
bool bind_frontier(tg_frontier *F, tg_graph *G) {
  if (!F->owner) {
    NEED(F->f.n == G->g.num_vertices(), "frontier size differs from the graph's vertex count");
    F->owner = G;
  }
  return F->owner == G;
}

void frontier_upload(tg_frontier *F, std::vector<uint32_t> &v, hipStream_t st) {
  F->f.queue.ensure(F->f.n, st);
  h2d(F->f.queue.get(), v.data(), v.size() * 4, st);
  F->f.count = v.size();
  F->f.all = false;
  F->f.has_queue = true;
  F->f.has_bits = false;
  TG_HIP(hipStreamSynchronize(st));
}
}  // namespace

int tg_frontier_set(tg_frontier *F, const uint32_t *verts, uint64_t count) {
  return guard([&] {
    NEED(F && (count == 0 || verts), "null argument");
    FrontierCtx c(F);
    uint32_t n = F->f.n;
    std::vector<uint32_t> v(verts, verts + count);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    NEED(v.empty() || v.back() < n, "vertex id out of range");
    frontier_upload(F, v, c.st);
  });
}

int tg_frontier_set_dense(tg_frontier *F, const uint8_t *els, uint64_t len) {
  return guard([&] {
    NEED(F && els, "null argument");
    NEED(len == F->f.n, "dense array length differs from the frontier size");
    FrontierCtx c(F);
    std::vector<uint32_t> v;
    for (uint64_t i = 0; i < len; i++)
      if (els[i]) v.push_back(static_cast<uint32_t>(i));
    frontier_upload(F, v, c.st);
  });
}

int tg_frontier_set_all(tg_frontier *F) {
  return guard([&] {
    NEED(F, "null argument");
    F->f.all = true;
    F->f.count = F->f.n;
    F->f.has_queue = false;
    F->f.has_bits = false;
  });
}

int tg_frontier_size(tg_frontier *F, uint64_t *out) {
  return guard([&] {
    NEED(F && out, "null argument");
    *out = F->f.count;
  });
}

int tg_frontier_get(tg_frontier *F, uint32_t *out, uint64_t cap, uint64_t *count) {
  return guard([&] {
    NEED(F && count, "null argument");
    FrontierCtx c(F);
    std::vector<uint32_t> v;
    if (F->owner) {
      F->owner->t.to_queue(F->f);
    } else if (F->f.all) {
      v.resize(F->f.n);
      for (uint32_t i = 0; i < F->f.n; i++) v[i] = i;
    }
    *count = F->f.count;
    if (!cap || !F->f.count) return;
    NEED(out, "null out");
    if (v.empty()) {  // the queue (an unbound frontier only ever holds a queue or "all")
      v.resize(F->f.count);
      d2h(v.data(), F->f.queue.get(), F->f.count * 4, c.st);
    }
    std::sort(v.begin(), v.end());
    std::copy(v.begin(), v.begin() + std::min<uint64_t>(cap, v.size()), out);
  });
}

int tg_edge_map(tg_graph *G, tg_frontier *in, int op, const tg_op_args *a, tg_frontier *out, uint32_t threshold) {
  return guard([&] {
    NEED(G && in && a, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    NEED(bind_frontier(in, G) && (!out || bind_frontier(out, G)), "frontier of another graph");
    if (op == TG_OP_TC) {
      // TC.h's countF: no output frontier (TC.h:78 passes output = false)
      NEED(a->count, "TC needs count");
      NEED(!out, "the TC op has no output frontier");
      if (in->f.all) {
        G->g.count_common_edges(nullptr, 0, reinterpret_cast<unsigned long long *>(a->count));
      } else {
        G->t.to_queue(in->f);
        G->g.count_common_edges(in->f.queue.get(), in->f.count, reinterpret_cast<unsigned long long *>(a->count));
      }
      return;
    }
    tg::OpState st;
    st.parent = a->parent;
    st.level = a->level;
    st.depth = a->depth;
    st.ids = a->ids;
    st.prev = a->prev;
    st.val = a->val;
    st.src_val = a->src_val;
    st.visited = a->visited;
    st.dist = a->dist;
    tg::Op o;
    switch (op) {
      case TG_OP_BFS: NEED(st.parent, "BFS needs parent"); o = tg::Op::BFS; break;
      case TG_OP_CC: NEED(st.ids, "CC needs ids"); o = tg::Op::CC; break;
      case TG_OP_BC_FWD: NEED(st.val && st.visited, "BC needs val and visited"); o = tg::Op::BC_FWD; break;
      case TG_OP_BC_BACK: NEED(st.val && st.visited, "BC needs val and visited"); o = tg::Op::BC_BACK; break;
      case TG_OP_PR: NEED(st.val && st.src_val, "PR needs val and src_val"); o = tg::Op::PR; break;
      case TG_OP_BF: NEED(st.dist && st.visited, "BF needs dist and visited"); o = tg::Op::BF; break;
      default: throw tg::Error(TG_EINVAL, "unknown op");
    }
    G->t.edge_map(in->f, o, st, out ? &out->f : nullptr, threshold);
    TG_HIP(hipStreamSynchronize(G->g.stream()));
  });
}

int tg_vertex_map(tg_graph *G, tg_frontier *f, int vop, const tg_vmap_args *a) {
  return guard([&] {
    NEED(f && a, "null argument");
    tg::Traversal::VArgs v;
    v.val = a->val;
    v.inv = a->inv;
    v.ids = a->ids;
    v.prev = a->prev;
    v.visited = a->visited;
    if (G) {
      std::lock_guard<std::mutex> l(G->g.mu);
      NEED(bind_frontier(f, G), "frontier of another graph");
    }
    tg_graph *owner = f->owner;
    if (!owner) {
      // no graph yet: the vertex ops other than PR_VERTEX read no graph state
      NEED(vop != TG_VOP_PR_VERTEX, "PR_VERTEX needs the graph (its degrees)");
      tg::vertex_map_unbound(f->f, vop, v, nullptr);
      TG_HIP(hipStreamSynchronize(nullptr));
      return;
    }
    std::lock_guard<std::mutex> l(owner->g.mu);
    owner->t.vertex_map(f->f, vop, v);
    TG_HIP(hipStreamSynchronize(owner->g.stream()));
  });
}

int tg_frontier_copy(tg_frontier *dst, tg_frontier *src) {
  return guard([&] {
    NEED(dst && src, "null argument");
    NEED(!dst->owner || !src->owner || dst->owner == src->owner, "frontiers of different graphs");
    NEED(dst->f.n == src->f.n, "frontiers of different sizes");
    tg_graph *owner = src->owner ? src->owner : dst->owner;
    if (!owner) {
      tg::copy_frontier_unbound(dst->f, src->f, nullptr);
      TG_HIP(hipStreamSynchronize(nullptr));
      return;
    }
    std::lock_guard<std::mutex> l(owner->g.mu);
    dst->owner = src->owner = owner;
    owner->t.copy_frontier(dst->f, src->f);
    TG_HIP(hipStreamSynchronize(owner->g.stream()));
  });
}

// ---------------------------------------------------------------- batching
int tg_bfs_multi(tg_graph *G, const uint32_t *srcs, uint32_t k, uint32_t *levels) {
  return guard([&] {
    NEED(G && srcs && levels, "null argument");
    NEED(k >= 1 && k <= 64, "k must be in [1, 64]");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    hipStream_t st = G->g.stream();
    G->s_u32.ensure(static_cast<uint64_t>(k) * n, st);
    G->t.bfs_multi(srcs, k, G->s_u32.get());
    d2h(levels, G->s_u32.get(), static_cast<uint64_t>(k) * n * 4, st);
  });
}

int tg_autotune(tg_graph *G, double *report) {
  return guard([&] {
    NEED(G && report, "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->t.autotune(report);
  });
}

int tg_get_tuning(tg_graph *G, uint32_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    const tg::Graph::Tuning &T = G->g.tune;
    out[0] = T.thr_bfs;
    out[1] = T.thr_ms;
    out[2] = T.thr_cc;
    out[3] = T.pull_block;
    out[4] = T.pr_graph;
    out[5] = T.pr_sliced;
    out[6] = T.push_wave;
    out[7] = T.level_graph;
  });
}

int tg_set_tuning(tg_graph *G, const uint32_t *in) {
  return guard([&] {
    NEED(G && in, "null argument");
    NEED(in[0] >= 1 && in[1] >= 1 && in[2] >= 1, "thresholds must be >= 1");
    NEED(in[3] == 64 || in[3] == 128 || in[3] == 256 || in[3] == 512, "pull block must be 64, 128, 256 or 512");
    NEED(in[4] <= 1 && in[5] <= 1 && in[7] <= 1, "pr_graph, pr_sliced and level_graph must be 0 or 1");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->g.tune = tg::Graph::Tuning{in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7]};
  });
}

int tg_bfs_multi_dev(tg_graph *G, const uint32_t *srcs, uint32_t k, uint32_t *levels) {
  return guard([&] {
    NEED(G && srcs && levels, "null argument");
    NEED(k >= 1 && k <= 64, "k must be in [1, 64]");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->t.bfs_multi(srcs, k, levels);
  });
}

int tg_bc_multi_dev(tg_graph *G, const uint32_t *srcs, uint32_t k, double *out) {
  return guard([&] {
    NEED(G && srcs && out, "null argument");
    NEED(k >= 1 && k <= 64, "k must be in [1, 64]");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->t.bc_multi(srcs, k, out);
  });
}

int tg_apply_ops(tg_graph *G, const uint32_t *src, const uint32_t *dst, const uint32_t *ops, uint64_t count) {
  return guard([&] {
    NEED(G && (count == 0 || (src && dst && ops)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    if (count == 0) return;
    for (uint64_t i = 0; i < count; i++) NEED(ops[i] <= 1, "op must be 0 (insert) or 1 (delete)");
    hipStream_t st = G->g.stream();
    G->s_src.ensure(count, st);
    G->s_dst.ensure(count, st);
    G->s_u32.ensure(count, st);
    TG_HIP(hipMemcpyAsync(G->s_src.get(), src, count * 4, hipMemcpyHostToDevice, st));
    TG_HIP(hipMemcpyAsync(G->s_dst.get(), dst, count * 4, hipMemcpyHostToDevice, st));
    TG_HIP(hipMemcpyAsync(G->s_u32.get(), ops, count * 4, hipMemcpyHostToDevice, st));
    G->g.apply_ops(G->s_src.get(), G->s_dst.get(), G->s_u32.get(), count);
    TG_HIP(hipStreamSynchronize(st));
  });
}

int tg_bc_multi(tg_graph *G, const uint32_t *srcs, uint32_t k, double *out) {
  return guard([&] {
    NEED(G && srcs && out, "null argument");
    NEED(k >= 1 && k <= 64, "k must be in [1, 64]");
    std::lock_guard<std::mutex> l(G->g.mu);
    uint32_t n = G->g.num_vertices();
    hipStream_t st = G->g.stream();
    G->s_f64.ensure(static_cast<uint64_t>(k) * n, st);
    G->t.bc_multi(srcs, k, G->s_f64.get());
    d2h(out, G->s_f64.get(), static_cast<uint64_t>(k) * n * 8, st);
  });
}

int tg_ltc(tg_graph *G, const uint32_t *verts, uint32_t q, uint64_t *counts) {
  return guard([&] {
    NEED(G && (q == 0 || (verts && counts)), "null argument");
    std::lock_guard<std::mutex> l(G->g.mu);
    G->t.ltc(verts, q, counts);
  });
}

int tg_tier_counts(tg_graph *G, uint64_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    std::lock_guard<std::mutex> lk(G->g.mu);
    G->g.tier_counts(out);
  });
}

}  // extern "C"

struct tg_server {
  tg_graph *owner;
  std::unique_ptr<tg::Server> s;
};

extern "C" {

int tg_server_create(tg_graph *G, int batched, uint64_t mem_budget, int keep_results, tg_server **out) {
  return guard([&] {
    NEED(G && out, "null argument");
    auto *sv = new tg_server;
    sv->owner = G;
    NEED(batched >= 0 && batched <= 2, "batched must be 0, 1 or 2");
    sv->s = std::make_unique<tg::Server>(G->g, batched, mem_budget, keep_results != 0);
    *out = sv;
  });
}

int tg_server_destroy(tg_server *s) {
  return guard([&] { delete s; });
}

int tg_server_submit(tg_server *S, const tg_request *r, uint64_t *id) {
  return guard([&] {
    NEED(S && r, "null argument");
    NEED(r->type >= TG_REQ_UPDATE && r->type <= TG_REQ_LTC, "unknown request type");
    uint32_t n = S->owner->g.num_vertices();
    tg::Request q;
    q.type = r->type;
    q.arrival = r->arrival_s;
    q.src = r->src;
    q.iters = r->iters;
    if (r->type == TG_REQ_UPDATE) {
      NEED((r->n_ins == 0 || (r->ins_src && r->ins_dst)) && (r->n_del == 0 || (r->del_src && r->del_dst)),
           "null update arrays");
      q.us.assign(r->ins_src, r->ins_src + r->n_ins);
      q.ud.assign(r->ins_dst, r->ins_dst + r->n_ins);
      q.uop.assign(r->n_ins, 0u);
      q.us.insert(q.us.end(), r->del_src, r->del_src + r->n_del);
      q.ud.insert(q.ud.end(), r->del_dst, r->del_dst + r->n_del);
      q.uop.insert(q.uop.end(), r->n_del, 1u);
      for (size_t i = 0; i < q.us.size(); i++) NEED(q.us[i] < n && q.ud[i] < n, "vertex id out of range");
    } else {
      NEED(r->type == TG_REQ_CC || r->type == TG_REQ_PR || r->src < n, "vertex id out of range");
    }
    uint64_t i = S->s->submit(std::move(q));
    if (id) *id = i;
  });
}

int tg_server_run(tg_server *S) {
  return guard([&] {
    NEED(S, "null server");
    // before the graph's lock, which a running loop holds
    NEED(!S->s->busy(), "the server's loop is already running");
    std::lock_guard<std::mutex> l(S->owner->g.mu);
    S->s->run();
  });
}

int tg_server_start(tg_server *S) {
  return guard([&] {
    NEED(S, "null server");
    S->s->start(&S->owner->g.mu);
  });
}

int tg_server_finish(tg_server *S) {
  return guard([&] {
    NEED(S, "null server");
    S->s->finish();
  });
}

int tg_server_wait(tg_server *S, uint64_t id, double timeout_s, int *status) {
  return guard([&] {
    NEED(S && status, "null argument");
    *status = S->s->wait(id, timeout_s);
  });
}

int tg_server_result(tg_server *S, uint64_t id, tg_result *out) {
  return guard([&] {
    NEED(S && out, "null argument");
    const int st = S->s->status(id);
    const tg::Request &r = S->s->result(id);
    *out = tg_result{};
    out->status = st;
    out->arrival_s = r.arrival;
    if (st == 1) return;  // pending: the other fields are not written yet
    out->start_s = r.start;
    out->done_s = r.done;
    out->version = r.version;
    out->digest = r.digest;
    out->value = r.value;
    out->fvalue = r.fvalue;
    out->group = r.group;
  });
}

int tg_server_result_data(tg_server *S, uint64_t id, void *out, uint64_t bytes) {
  return guard([&] {
    NEED(S && out, "null argument");
    NEED(S->s->status(id) != 1, "request still pending");
    const tg::Request &r = S->s->result(id);
    if (!r.r32.empty()) {
      NEED(bytes >= r.r32.size() * 4, "buffer too small");
      std::memcpy(out, r.r32.data(), r.r32.size() * 4);
    } else if (!r.r64.empty()) {
      NEED(bytes >= r.r64.size() * 8, "buffer too small");
      std::memcpy(out, r.r64.data(), r.r64.size() * 8);
    } else {
      throw tg::Error(TG_EINVAL, "no kept result for this request");
    }
  });
}

int tg_server_summary(tg_server *S, double *out) {
  return guard([&] {
    NEED(S && out, "null argument");
    std::vector<double> v = S->s->summary();
    for (size_t i = 0; i < v.size(); i++) out[i] = v[i];
  });
}

// ---------------------------------------------------------------- helpers
int tg_rmat_dev(uint32_t n, uint32_t seed, double a, double b, double c, int nominal, uint64_t start,
                uint64_t count, int flags, uint32_t *src, uint32_t *dst, uint64_t *written) {
  return guard([&] {
    NEED(count == 0 || (src && dst), "null argument");
    uint64_t w = tg::rmat_device(n, seed, a, b, c, nominal, start, count, flags, src, dst, nullptr);
    if (written) *written = w;
  });
}

int tg_dev_alloc(uint64_t bytes, void **out) {
  return guard([&] {
    NEED(out, "null out");
    *out = nullptr;
    TG_HIP(hipMalloc(out, bytes ? bytes : 1));
  });
}

int tg_dev_free(void *p) {
  return guard([&] {
    if (p) TG_HIP(hipFree(p));
  });
}

int tg_memcpy_h2d(void *dst, const void *src, uint64_t bytes) {
  return guard([&] { TG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}

int tg_memcpy_d2h(void *dst, const void *src, uint64_t bytes) {
  return guard([&] { TG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

int tg_memcpy_d2d(void *dst, const void *src, uint64_t bytes) {
  return guard([&] {
    tg::copy_bytes(dst, src, bytes, nullptr);
    TG_HIP(hipStreamSynchronize(nullptr));
  });
}

int tg_read_adjacency_graph(const char *path, uint32_t *n, uint64_t *m, uint32_t *src, uint32_t *dst, uint64_t cap) {
  return guard([&] {
    NEED(path && n && m, "null argument");
    tg::AdjFile f = tg::read_adjacency_graph(path);
    *n = f.n;
    *m = f.m;
    if (src || dst) {
      NEED(src && dst, "src and dst must both be given");
      NEED(cap >= f.m, "capacity below m");
      std::copy(f.src.begin(), f.src.end(), src);
      std::copy(f.dst.begin(), f.dst.end(), dst);
    }
  });
}

int tg_load_adjacency_graph(const char *path, tg_graph **out) {
  return guard([&] {
    NEED(path && out, "null argument");
    *out = nullptr;
    tg::AdjFile f = tg::read_adjacency_graph(path);
    std::unique_ptr<tg_graph> G(new tg_graph(f.n));
    if (f.m) {
      stage_pair(G.get(), f.src.data(), f.dst.data(), f.m);
      G->g.insert_batch(G->s_src.get(), G->s_dst.get(), f.m);
    }
    *out = G.release();
  });
}

int tg_device_sync(void) {
  return guard([&] { TG_HIP(hipDeviceSynchronize()); });
}

int tg_profile_enable(tg_graph *G, int on) {
  return guard([&] {
    NEED(G, "null graph");
    G->t.profiling = on != 0;
  });
}

int tg_profile_reset(tg_graph *G) {
  return guard([&] {
    NEED(G, "null graph");
    G->t.ktimes.clear();
    if (G->span_a) (void)hipEventDestroy(G->span_a);
    G->span_a = nullptr;
    G->span_steps = 0;
  });
}

int tg_profile_get(tg_graph *G, const char *name, double *ms, uint64_t *launches) {
  return guard([&] {
    NEED(G && name && ms && launches, "null argument");
    *ms = 0;
    *launches = 0;
    for (auto &k : G->t.ktimes)
      if (k.name == name) {
        *ms = k.ms;
        *launches = k.launches;
      }
    if (std::strcmp(name, "update_step") == 0 && G->span_a && G->span_steps) {
      float span = 0;
      if (!G->span_b) TG_HIP(hipEventCreate(&G->span_b));
      TG_HIP(hipEventRecord(G->span_b, G->g.stream()));
      TG_HIP(hipEventSynchronize(G->span_b));
      TG_HIP(hipEventElapsedTime(&span, G->span_a, G->span_b));
      *ms = span;
      *launches = G->span_steps;
    }
  });
}

int tg_debug_fail_next(tg_graph *G, int what) {
  return guard([&] {
    NEED(G && (what == 1 || what == 2), "what must be 1 or 2");
    std::lock_guard<std::mutex> lk(G->g.mu);
    G->g.debug_fail |= static_cast<uint32_t>(what);
  });
}

int tg_stats(tg_graph *G, uint64_t *out) {
  return guard([&] {
    NEED(G && out, "null argument");
    const tg::Stats &s = G->g.stats();
    out[0] = s.pma_rebuilds;
    out[1] = s.pma_windows;
    out[2] = s.pma_window_leaves;
    out[3] = s.tree_blocks_written;
    out[4] = s.promotions;
    out[5] = G->g.num_trees();
    out[6] = G->g.pma_slots();
    out[7] = s.tree_small_updates;
    out[8] = s.pma_leaf_merges;
    out[9] = G->g.pma_elements();
    out[10] = s.last_ins_keys;
    out[11] = s.last_ins_sources;
    out[12] = s.last_del_keys;
    out[13] = s.last_del_sources;
    out[14] = s.pma_low_windows;
    out[15] = s.pma_shrinks;
    out[16] = s.sum_ins_sources;
    out[17] = s.sum_del_sources;
  });
}

int tg_pma_leaf_counts(tg_graph *G, uint32_t *out, uint64_t cap, uint64_t *leaves) {
  return guard([&] {
    NEED(G && leaves && (out || cap == 0), "null argument");
    *leaves = G->g.pma_leaf_counts(out, cap);
  });
}

}  // extern "C"
