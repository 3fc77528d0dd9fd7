/*
 * terrace_hip.h — C ABI of the MI355X (gfx950) Terrace hot path.
 *
 * A drop-in for the batch-update + traversal path of the reference's
 * header-only C++ container (namespace graphstore, include/graph.h) and its
 * Ligra-style traversal (include/Map.cpp, include/VertexSubset.cpp) plus the
 * algorithm entry points the benchmark driver calls (src/graph_bm.cc).
 *
 * Conventions
 *   - Every function returns an int status: TG_OK (0) or a negative code.
 *     Nothing aborts and no C++ exception crosses the ABI; tg_last_error()
 *     returns the message of the last failure on the calling thread.
 *   - Plain pointers and sizes only.  Functions ending in _dev take device
 *     pointers (HBM); all others take host pointers.
 *   - Calls on one tg_graph are serialised internally; each graph owns one
 *     HIP stream (default priority).
 *   - Vertex ids are uint32 in [0, n); edges are directed (src, dst).
 */
#ifndef TERRACE_HIP_H
#define TERRACE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TG_OK 0
#define TG_EINVAL (-1)   /* bad argument (id out of range, null pointer) */
#define TG_ENOMEM (-2)   /* device allocation failed before the call changed anything;
                            the graph is unchanged */
#define TG_EDEVICE (-3)  /* HIP error; message via tg_last_error() */
/* Internal check failed.  When an update batch fails after it began changing
 * the graph (an allocation for the graph's own growth: PMA, B-tree blocks or
 * index; or an internal consistency check), the adjacency may be
 * half-updated: the graph is marked invalid, the call returns TG_EINTERNAL,
 * and every later call on that graph returns TG_EINTERNAL.  A failure in the
 * batch's preparation (sort, existence check, source runs) is reported
 * before any write and leaves the graph valid and unchanged. */
#define TG_EINTERNAL (-4)

typedef struct tg_graph tg_graph;
typedef struct tg_frontier tg_frontier;

const char *tg_last_error(void);

/* ---- container lifetime: Graph(uint32_t) / ~Graph()
 *      include/graph.h:118-120, 232-239 ------------------------------- */
int tg_create(uint32_t num_vertices, tg_graph **out);
int tg_destroy(tg_graph *g);
/* ---- weighted graphs: the reference's WEIGHTED build (include/graph.h:
 * 34-38, 76-86, add_edge_batch with weights :430-436).  One uint32 weight
 * below 2^31 per directed edge.  A batch keeps the first weight given for
 * an edge; inserting an existing edge leaves its weight; deletes drop it.
 * On a weighted graph, tg_insert_batch, tg_add_edge and the coalesced ops
 * return TG_EINVAL (they carry no weights); tg_sssp_bf uses the weights as
 * edge lengths (include/BellmanFord.h:67-143). */
int tg_create_weighted(uint32_t num_vertices, tg_graph **out);
int tg_is_weighted(tg_graph *g, int *out);
int tg_insert_batch_weighted(tg_graph *g, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                             uint64_t count);
int tg_insert_batch_weighted_dev(tg_graph *g, const uint32_t *src, const uint32_t *dst, const uint32_t *w,
                                 uint64_t count);
/* Graph::is_edge of the weighted build returns the weight: out[i] = weight
 * of (src[i], dst[i]), or 0xFFFFFFFF when the edge is absent. */
int tg_edge_weights(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t count, uint32_t *out);

/* ---- batch updates ---------------------------------------------------
 * tg_insert_batch replaces Graph::add_edge_batch (include/graph.h:434-502).
 * Unlike the reference it accepts any order and duplicates (it sorts and
 * deduplicates on the device), does not modify its inputs and needs no perm
 * vector.  Edges already present are ignored.  Tier routing is the
 * reference's: the 13 smallest neighbours inline, the rest in the PMA while
 * degree + run length <= 1024, else in the vertex's B-tree.
 * tg_delete_batch is the batched Graph::remove_edge (include/graph.h:
 * 675-791); deleting an absent edge is a no-op. */
int tg_insert_batch(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t count);
int tg_delete_batch(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t count);
int tg_insert_batch_dev(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t count);
int tg_delete_batch_dev(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t count);
/* A mixed update batch: add_edge_batch of the n_ins inserts followed by
 * remove_edge of the n_del deletes, as the benchmark driver applies a batch
 * (src/graph_bm.cc:354-378), in one call that sorts and prepares both
 * halves together.  The result equals tg_insert_batch then tg_delete_batch:
 * an edge both inserted and deleted ends absent; absent deletes are no-ops.
 * n_ins + n_del < 2^31. */
int tg_update_batch(tg_graph *g, const uint32_t *ins_src, const uint32_t *ins_dst, uint64_t n_ins,
                    const uint32_t *del_src, const uint32_t *del_dst, uint64_t n_del);
int tg_update_batch_dev(tg_graph *g, const uint32_t *ins_src, const uint32_t *ins_dst, uint64_t n_ins,
                        const uint32_t *del_src, const uint32_t *del_dst, uint64_t n_del);
/* Graph::add_edge / remove_edge (include/graph.h:507-641, 675-791):
 * *result = 0 when the edge was added/removed, -1 when it was already
 * present/absent. */
int tg_add_edge(tg_graph *g, uint32_t s, uint32_t d, int *result);
int tg_remove_edge(tg_graph *g, uint32_t s, uint32_t d, int *result);

/* ---- queries ---------------------------------------------------------- */
/* Graph::is_edge (include/graph.h:643-673): *out = 1 or 0 */
int tg_is_edge(tg_graph *g, uint32_t s, uint32_t d, uint32_t *out);
int tg_is_edge_batch(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t count, uint8_t *out);
/* Graph::count_common (include/graph.h:1085-1255, declared :171-211) for q
 * pairs at once: out[i] = |N(src[i]) ∩ N(dst[i]) ∩ [0, min(src[i], dst[i]))|,
 * the common neighbours below both endpoints, over every pair of tiers
 * (inline, PMA, B-tree).  _dev: device arrays. */
int tg_count_common(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out);
int tg_count_common_dev(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out);
/* tg_count_common without any device allocation; no counterpart in the
 * reference.  For callers that cannot allocate device memory (serving
 * loops, code under HIP graph capture, harnesses that forbid hipMalloc):
 * host arrays are staged through a static device array of 12 MiB that each
 * GPU holds until the process ends, so q is at most 2^20 pairs (larger q:
 * TG_EINVAL).  Same results and errors as tg_count_common.  Calls on one
 * GPU share the array: the library serialises them within the process. */
int tg_count_common_noalloc(tg_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out);
/* Graph::degree (include/graph.h:797-803) */
int tg_degree(tg_graph *g, uint32_t v, uint32_t *out);
int tg_degrees_all(tg_graph *g, uint32_t *out /* [n] */);
/* Graph::get_num_edges (include/graph.h:817-822), as uint64 */
int tg_num_edges(tg_graph *g, uint64_t *out);
/* Graph::get_num_vertices (include/graph.h:824-826) */
int tg_num_vertices(tg_graph *g, uint32_t *out);
/* Graph::get_size (include/graph.h:805-815): device bytes held */
int tg_memory_bytes(tg_graph *g, uint64_t *out);
/* NeighborIterator (include/graph.h:844-899) for every vertex at once:
 * offsets[n+1], nbrs[num_edges], each list in sorted order. */
int tg_export_adjacency(tg_graph *g, uint64_t *offsets, uint32_t *nbrs);
int tg_export_adjacency_dev(tg_graph *g, uint64_t *offsets, uint32_t *nbrs);
/* Graph::neighbors(v): copies min(degree, cap) neighbours, sorted. */
int tg_neighbors(tg_graph *g, uint32_t v, uint32_t *out, uint32_t cap, uint32_t *degree);
/* Tier of every vertex: 0 inline only, 1 PMA, 2 B-tree. */
int tg_tiers(tg_graph *g, uint8_t *out /* [n] */);
/* Tier sizes: out[0] PMA-tier vertices, out[1] their edges beyond the 13
 * inline slots, out[2] B-tree vertices, out[3] their edges beyond the inline
 * slots.  Graph::get_size's breakdown (include/graph.h:805-815) by tier. */
int tg_tier_counts(tg_graph *g, uint64_t *out /* [4] */);

/* ---- algorithms (run_algorithm's tests, src/graph_bm.cc:86-229) ------- */
/* BFS_with_edge_map (include/BFS.h:41-84): parents[n] (-1 unreached),
 * levels[n] (UINT32_MAX unreached); either may be NULL. */
int tg_bfs(tg_graph *g, uint32_t src, int32_t *parents, uint32_t *levels);
/* PR_S<double> (include/Pagerank.h:106-184), undamped as in the reference */
int tg_pagerank(tg_graph *g, uint32_t iters, double *out);
/* CC (include/Components.h:92-120): minimum vertex id per component */
int tg_cc(tg_graph *g, uint32_t *labels);
/* BC (include/BC.h:99-158): unreachable vertices are NaN as in the reference */
int tg_bc(tg_graph *g, uint32_t src, double *out);
/* TC (include/TC.h:73-87) */
int tg_tc(tg_graph *g, uint64_t *count);
/* SSSP_BF (include/BellmanFordUnweighted.h:60-130, graph_bm.cc test
 * "SSSP_BF"): dist[n] = hop distance from src, INT_MAX/2 when unreachable.
 * On a weighted graph: BellmanFord.h's distances with the stored weights as
 * lengths (sums saturate at INT_MAX/2). */
int tg_sssp_bf(tg_graph *g, uint32_t src, int32_t *dist);
/* Device-output variants (pointers into HBM) */
int tg_bfs_dev(tg_graph *g, uint32_t src, int32_t *parents, uint32_t *levels, uint32_t *depth);
int tg_pagerank_dev(tg_graph *g, uint32_t iters, double *out);
int tg_cc_dev(tg_graph *g, uint32_t *labels, uint32_t *rounds);
int tg_bc_dev(tg_graph *g, uint32_t src, double *out);
int tg_sssp_bf_dev(tg_graph *g, uint32_t src, int32_t *dist, uint32_t *rounds);

/* ---- frontiers and edgeMap (include/VertexSubset.cpp, include/Map.cpp) */
enum tg_op { TG_OP_BFS = 1, TG_OP_CC = 2, TG_OP_BC_FWD = 3, TG_OP_BC_BACK = 4, TG_OP_PR = 5, TG_OP_TC = 6,
             TG_OP_BF = 7 };
/* Device arrays an op reads and writes (all device pointers):
 *   BFS     parent (int32, -1 = unvisited), level (optional), depth
 *   CC      ids, prev (prevIDs; output = vertices whose id changed)
 *   BC_*    val (NumPaths / Dependencies), visited (uint8)
 *   PR      src_val (normalised p_curr), val (p_next, assigned)
 *   BF      dist (int32 ShortestPathLen), visited (claimed: the array is
 *           written and must be padded to a multiple of 4 bytes)
 *   TC      count (uint64 device counter): TC.h's countF (TC.h:51-70);
 *           count += count_common(u, v) for every frontier vertex u and
 *           neighbour v < u; no output frontier (out must be NULL) */
typedef struct {
  int32_t *parent;
  uint32_t *level;
  uint32_t depth;
  uint32_t *ids;
  const uint32_t *prev;
  double *val;
  const double *src_val;
  const uint8_t *visited;
  int32_t *dist;
  uint64_t *count;
} tg_op_args;
/* An empty frontier of graph g's vertex count. */
int tg_frontier_create(tg_graph *g, tg_frontier **out);
/* VertexSubset(el_t e, uint64_t max_el, bool all) and VertexSubset(bool *els,
 * uint64_t len) (include/VertexSubset.cpp:109-128) carry no graph: an empty
 * frontier over [0, n) that the first tg_edge_map (or tg_vertex_map with a
 * graph) binds to its graph, whose vertex count must be n.  Until then
 * tg_vertex_map accepts g = NULL for every vertex op but PR_VERTEX. */
int tg_frontier_create_n(uint32_t n, tg_frontier **out);
/* members = { i : els[i] != 0 }, len = the frontier's n */
int tg_frontier_set_dense(tg_frontier *f, const uint8_t *els, uint64_t len);
int tg_frontier_destroy(tg_frontier *f);
int tg_frontier_set(tg_frontier *f, const uint32_t *verts, uint64_t count); /* host list */
int tg_frontier_set_all(tg_frontier *f);
int tg_frontier_size(tg_frontier *f, uint64_t *out);                       /* get_n */
/* sorted members into out[cap] (host); *count = size */
int tg_frontier_get(tg_frontier *f, uint32_t *out, uint64_t cap, uint64_t *count);
/* edgeMap(G, vs, F, output, threshold) (include/Map.cpp:66-74); out may be
 * NULL when output is false. */
int tg_edge_map(tg_graph *g, tg_frontier *in, int op, const tg_op_args *args, tg_frontier *out,
                uint32_t threshold);

/* vertexMap(vs, F, output=false) (include/Map.cpp:76-108) for the vertex
 * functors of the algorithm headers, applied to the members of f:
 *   PR_VERTEX      val[i] = val[i] / degree(i)        (Pagerank.h:61-69)
 *   PR_RESET       val[i] = 0                         (Pagerank.h:72-81)
 *   CC_VERTEX      prev[i] = ids[i]                   (Components.h:37-43)
 *   BC_VERTEX      visited[i] = 1                     (BC.h:74-81)
 *   BC_BACK_VERTEX visited[i] = 1; val[i] += inv[i]   (BC.h:84-94)
 * and the parallel_for loops of BC (BC.h:126-128, 130-132, 140-142):
 *   INVERT         val[i] = 1 / val[i]
 *   CLEAR_VISITED  visited[i] = 0
 *   BC_FINAL       val[i] = (val[i] - inv[i]) / inv[i] */
enum tg_vop { TG_VOP_PR_VERTEX = 1, TG_VOP_PR_RESET = 2, TG_VOP_CC_VERTEX = 3, TG_VOP_BC_VERTEX = 4,
              TG_VOP_BC_BACK_VERTEX = 5, TG_VOP_INVERT = 6, TG_VOP_CLEAR_VISITED = 7, TG_VOP_BC_FINAL = 8 };
typedef struct {
  double *val;
  const double *inv;
  const uint32_t *ids;
  uint32_t *prev;
  uint8_t *visited;
} tg_vmap_args;
int tg_vertex_map(tg_graph *g, tg_frontier *f, int vop, const tg_vmap_args *args); /* g: see create_n */
/* Copy of a frontier (VertexSubset(other, true)). */
int tg_frontier_copy(tg_frontier *dst, tg_frontier *src);

/* ---- multi-source BFS and local triangle counts -----------------------
 * No counterpart in the reference (serving / batching layer).
 * tg_bfs_multi: k <= 64 BFS from srcs[] in one traversal (one bit per
 * source); levels[j*n + v] = level of v from srcs[j] (UINT32_MAX
 * unreached), identical to tg_bfs's levels for each source (host arrays).
 * tg_ltc: triangles through each of verts[0..q) (host arrays). */
int tg_bfs_multi(tg_graph *g, const uint32_t *srcs, uint32_t k, uint32_t *levels);
/* k <= 64 single-source BC runs (BC.h:99-158) in one traversal:
 * out[j*n + v] = BC of v from srcs[j], NaN where unreachable. */
int tg_bc_multi(tg_graph *g, const uint32_t *srcs, uint32_t k, double *out);
/* Device-output forms: levels / out are device arrays of k*n entries. */
int tg_bfs_multi_dev(tg_graph *g, const uint32_t *srcs, uint32_t k, uint32_t *levels);
int tg_bc_multi_dev(tg_graph *g, const uint32_t *srcs, uint32_t k, double *out);
/* Coalesced updates: ops[i] = 0 insert / 1 delete of (src[i], dst[i]), in
 * order; one device pass whose result equals applying them one by one (the
 * last op on an edge wins).  Host arrays.  No counterpart in the reference
 * (its drivers call add_edge_batch / remove_edge per batch,
 * src/graph_bm.cc:363-378). */
int tg_apply_ops(tg_graph *g, const uint32_t *src, const uint32_t *dst, const uint32_t *ops, uint64_t count);

/* ---- autotuner ---------------------------------------------------------
 * Sweeps the edgeMap push/pull threshold (dense when n / t <= |frontier|,
 * include/Map.cpp:66-74, reference t = 20) of BFS, multi-source BFS and CC
 * and the block size of the inline / PMA pull kernels on the current graph;
 * keeps the fastest for later whole-algorithm calls on this handle.
 * report[12]: thr_bfs, thr_ms, thr_cc, pull_block, then (reference ms,
 * tuned ms) for BFS (3 sources), MS-BFS (64 sources), CC and PageRank (3
 * iterations).  Results never depend on the tuning.
 * tuning[8]: thr_bfs, thr_ms, thr_cc, pull_block (64/128/256/512), pr_graph
 * (1 = PageRank iterations replayed from a captured HIP graph, kept per
 * graph version), pr_sliced (1 = PageRank over an XCD-sliced adjacency copy
 * when n <= 6M: each XCD's L2 holds one slice of the ranks), push_wave (push
 * frontiers up to this size give each PMA vertex a 64-lane wave, larger
 * ones 16 lanes), level_graph (1 = BFS and CC levels, whose directions are
 * chosen on the device, replayed from a captured HIP graph of two levels;
 * 0 = the same kernels launched one by one). */
int tg_autotune(tg_graph *g, double *report);
int tg_get_tuning(tg_graph *g, uint32_t *tuning);
int tg_set_tuning(tg_graph *g, const uint32_t *tuning);
int tg_ltc(tg_graph *g, const uint32_t *verts, uint32_t q, uint64_t *counts);

/* ---- request loop (serving) --------------------------------------------
 * No counterpart in the reference.  A server runs requests at their arrival
 * offsets (seconds from tg_server_run / tg_server_start) against one graph.
 * tg_server_run replays the requests submitted before it and returns when
 * they are done.  tg_server_start runs the same loop on a thread of its own
 * and returns at once: tg_server_submit stays open (from any thread), a
 * request submitted while the loop runs arrives at its submission time (or
 * at arrival_s, if later), tg_server_wait waits for one request, and
 * tg_server_finish closes the input, waits for every request and returns
 * the loop's status.  The loop holds the graph while it runs: other calls
 * on that graph wait until tg_server_finish.
 * batched = 0: one request at a time in arrival order (the plain path).
 * batched = 1: continuous batching; every query still runs on the version
 * that contains exactly the updates that arrived before it.
 * batched = 2: continuous batching with snapshot-at-admission; updates are
 * applied as they arrive, queries are pooled and admitted together on the
 * version current at admission (it contains every update that arrived
 * before the query, and may contain updates that arrived while it waited).
 * tg_result.version names the version; the result equals the query run
 * alone on it.  Requests whose footprint exceeds mem_budget bytes are
 * rejected (status TG_ENOMEM). */
typedef struct tg_server tg_server;
enum tg_req { TG_REQ_UPDATE = 1, TG_REQ_BFS = 2, TG_REQ_CC = 3, TG_REQ_PR = 4, TG_REQ_LTC = 5 };
typedef struct {
  int type;
  double arrival_s;
  uint32_t src;    /* BFS source / LTC vertex */
  uint32_t iters;  /* PR iterations */
  const uint32_t *ins_src, *ins_dst;
  uint64_t n_ins;
  const uint32_t *del_src, *del_dst;
  uint64_t n_del;
} tg_request;
typedef struct {
  int status; /* 0 done, 1 pending, < 0 error code */
  double arrival_s, start_s, done_s;
  uint64_t version; /* update batches applied before it ran */
  uint64_t digest;  /* order-independent digest of levels / labels / ranks */
  uint64_t value;   /* LTC count */
  double fvalue;    /* PR: sum of ranks */
  uint64_t group;
} tg_result;
int tg_server_create(tg_graph *g, int batched, uint64_t mem_budget, int keep_results, tg_server **out);
int tg_server_destroy(tg_server *s);
int tg_server_submit(tg_server *s, const tg_request *r, uint64_t *id);
int tg_server_run(tg_server *s);
int tg_server_start(tg_server *s);
int tg_server_finish(tg_server *s);
/* *status = the request's status after waiting up to timeout_s seconds for
 * it to complete (1: still pending, or the loop is not running) */
int tg_server_wait(tg_server *s, uint64_t id, double timeout_s, int *status);
/* while a request is pending only status and arrival_s are filled in */
int tg_server_result(tg_server *s, uint64_t id, tg_result *out);
/* kept result array (keep_results=1): BFS levels / CC labels as uint32[n],
 * PR ranks as double[n] */
int tg_server_result_data(tg_server *s, uint64_t id, void *out, uint64_t bytes);
/* [0] makespan s, [1] requests/s, [2] p50 s, [3] p99 s, [4] p50 update,
 * [5] p99 update, [6] p50 query, [7] p99 query, [8] groups, [9] rejected
 * (over the completed requests) */
int tg_server_summary(tg_server *s, double *out /* [10] */);

/* ---- workload and device memory helpers -------------------------------
 * rMat<uint32_t>(n, seed, a, b, c)(i) for i in [start, start+count)
 * (include/rmat_util.h:81-115) written straight into device arrays.
 * nominal=0 is the reference's divisor (INT32_MAX), 1 divides by 2^32.
 * flags: 1 = also write every reverse edge, 2 = drop self-loops.
 * *written = entries written (output order is not index order). */
int tg_rmat_dev(uint32_t n, uint32_t seed, double a, double b, double c, int nominal, uint64_t start,
                uint64_t count, int flags, uint32_t *src, uint32_t *dst, uint64_t *written);
int tg_dev_alloc(uint64_t bytes, void **out);
int tg_dev_free(void *p);
int tg_memcpy_h2d(void *dst, const void *src, uint64_t bytes);
int tg_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
/* Ligra AdjacencyGraph text file (get_edges_from_file_adj_sym,
 * include/io_util.h:112-145): "AdjacencyGraph", n, m, n offsets, m ids.
 * tg_read_adjacency_graph fills n and m, and with src/dst given (room for
 * cap >= m entries) the m edges in file order.  A malformed file returns
 * TG_EINVAL with the reason in tg_last_error (the reference exits).
 * tg_load_adjacency_graph creates a graph of n vertices and inserts the
 * file's edges as one batch (the load in src/graph_bm.cc:236-266). */
int tg_read_adjacency_graph(const char *path, uint32_t *n, uint64_t *m, uint32_t *src, uint32_t *dst,
                            uint64_t cap);
int tg_load_adjacency_graph(const char *path, tg_graph **out);
/* Device-to-device copy by a plain vectorised kernel (synchronous). */
int tg_memcpy_d2d(void *dst, const void *src, uint64_t bytes);
int tg_device_sync(void);

/* ---- measurement -------------------------------------------------------
 * With profiling on, kernels of interest are bracketed by hipEvents on
 * the graph's stream; tg_profile_get sums their durations by name.
 * "update_step" is the span of tg_update_batch_dev calls since the last
 * reset: an event before the first call, and one that tg_profile_get
 * records on the graph's stream, with launches = the calls (nothing is
 * recorded between calls). */
int tg_profile_enable(tg_graph *g, int on);
int tg_profile_reset(tg_graph *g);
int tg_profile_get(tg_graph *g, const char *name, double *ms, uint64_t *launches);
/* internal counters: [0] PMA rebuilds, [1] PMA windows, [2] window leaves,
 * [3] tree blocks written, [4] promotions, [5] trees, [6] PMA slots,
 * [7] tree updates rewritten in place (no block split or emptied),
 * [8] PMA inserts merged inside their leaves (no window redistribution),
 * [9] elements held in the PMA; the last update batch: [10] unique
 * inserted edges (present ones included), [11] their distinct sources,
 * [12] deleted edges that existed, [13] the distinct sources of the
 * deleted edges (of those that existed, for batches of more than 2^18
 * edges); [14] underfull PMA windows respread after deletes, [15] PMA
 * arrays shrunk after deletes; since the graph was created: [16] the sum
 * of [11] and [17] the sum of [13] over the update batches */
int tg_stats(tg_graph *g, uint64_t *out /* [18] */);
/* Element count of every PMA leaf (64 slots each), the per-leaf counts
 * that the density bounds of src/PMA.cc:382-394 are checked on (the
 * reference's check_no_full_leaves, PMA.cc, is the same kind of
 * diagnostic).  Writes min(cap, leaves) counts; *leaves gets the number of
 * leaves. */
int tg_pma_leaf_counts(tg_graph *g, uint32_t *out, uint64_t cap, uint64_t *leaves);
/* Failure injection for tests of the error paths; no counterpart in the
 * reference.  what = 1: the next update batch's preparation reports a
 * look-back error (TG_EINTERNAL, graph unchanged and still valid); what = 2:
 * the next batch fails as if an allocation failed after its first write
 * (TG_EINTERNAL, graph marked invalid).  Every kernel still runs to its
 * normal end. */
int tg_debug_fail_next(tg_graph *g, int what);

#ifdef __cplusplus
}
#endif
#endif
