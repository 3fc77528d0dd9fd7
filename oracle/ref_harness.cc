// Reference-build driver — TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile against the reference's own headers and
// sources where they lie (graph.h, PMA.cc, BFS.h, ...), into oracle/_ref/.
// It exists to produce golden vectors (tests/golden/make_golden.py) and to
// time the reference's CPU insert path (bench.py cpu_baseline,
// kind "reference").  It mirrors run_algorithm in src/graph_bm.cc:231-391:
// static build with add_edge_batch, insert batches, then BFS / PR / CC /
// BC / TC / SSSP_BF through the reference's own algorithm headers.
//
// Usage:
//   ref_harness rmat  <n> <seed> <a> <b> <c> <count> <out.bin>
//       writes count (u32 src, u32 dst) pairs of rMat<uint32_t>(i)
//   ref_harness run   <graph.bin> <out_prefix> <src> <iters> [batch.bin...]
//       graph.bin / batch.bin: u64 n, u64 m, then m (u32 src, u32 dst) pairs
//       sorted by (src,dst), unique.  Writes <out_prefix>.adj<k> and
//       .btree<k> (B-tree tier membership) after the static build (k=0) and
//       after every batch (k=1..), then .bfs .pr .cc .bc .tc .sssp for the
//       final graph.
//   ref_harness bench <graph.bin> <batch_size> <batches> <seed>
//       static build, then `batches` rMAT insert batches of batch_size
//       edges (sorted + unique, both directions), prints JSON timing.
//   ref_harness mixed <graph.bin> <ins0.bin> <del0.bin> [<ins1.bin> <del1.bin> ...]
//       static build, then per batch add_edge_batch(ins) and a parallel
//       remove_edge loop over del; prints JSON timing.

#define ENABLE_LOCK 1
#define WEIGHTED 0
#define VERIFY 0

#include <stdlib.h>
#include <sys/time.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "graph.h"
#include "parallel.h"
#include "util.h"
#include "rmat_util.h"
#include "BFS.h"
#include "Pagerank.h"
#include "Components.h"
#include "BC.h"
#include "TC.h"
#include "BellmanFordUnweighted.h"

using namespace graphstore;

namespace {

struct EdgeFile {
  uint64_t n = 0;
  std::vector<uint32_t> src, dst;
};

bool read_edges(const char *path, EdgeFile &ef) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  uint64_t hdr[2];
  if (fread(hdr, sizeof(uint64_t), 2, f) != 2) { fclose(f); return false; }
  ef.n = hdr[0];
  std::vector<uint32_t> pairs(2 * hdr[1]);
  if (hdr[1] && fread(pairs.data(), sizeof(uint32_t), 2 * hdr[1], f) != 2 * hdr[1]) {
    fclose(f);
    return false;
  }
  fclose(f);
  ef.src.resize(hdr[1]);
  ef.dst.resize(hdr[1]);
  for (uint64_t i = 0; i < hdr[1]; i++) {
    ef.src[i] = pairs[2 * i];
    ef.dst[i] = pairs[2 * i + 1];
  }
  return true;
}

void write_blob(const std::string &path, const void *p, size_t bytes) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(2); }
  if (bytes) fwrite(p, 1, bytes, f);
  fclose(f);
}

// add_edge_batch mutates dests (graph.h:346) and requires >=1 edge.
void insert(Graph &g, const EdgeFile &ef) {
  if (ef.src.empty()) return;
  std::vector<uint32_t> s(ef.src), d(ef.dst);
  std::vector<uint32_t> perm(s.size());
  for (uint32_t i = 0; i < perm.size(); i++) perm[i] = i;
  g.add_edge_batch(s.data(), d.data(), (uint32_t)s.size(), perm);
}

// Graph::is_btree (graph.h:192) is private.  An explicit template
// instantiation may name a private member (access checks do not apply
// there), so this stores a pointer to it that the dump below calls; the
// reference's sources are compiled unchanged.
struct IsBtreeTag {
  typedef bool (Graph::*type)(const Graph::vertex) const;
  friend type is_btree_ptr(IsBtreeTag);
};
template <typename Tag, typename Tag::type M>
struct ExposeMember {
  friend typename Tag::type is_btree_ptr(Tag) { return M; }
};
template struct ExposeMember<IsBtreeTag, &Graph::is_btree>;

// B-tree tier membership: u8 per vertex, 1 when aux_neighbors is set.
void dump_btree(Graph &g, uint64_t n, const std::string &path) {
  std::vector<uint8_t> out(n);
  for (uint64_t v = 0; v < n; v++) out[v] = (g.*is_btree_ptr(IsBtreeTag{}))((uint32_t)v) ? 1 : 0;
  write_blob(path, out.data(), out.size());
}

// Adjacency dump: u32 degree[n], then all neighbours in NeighborIterator
// order (inline, then PMA or B-tree in order).
void dump_adjacency(Graph &g, uint64_t n, const std::string &path) {
  std::vector<uint32_t> out(n);
  std::vector<uint32_t> nb;
  for (uint64_t v = 0; v < n; v++) {
    out[v] = g.degree((uint32_t)v);
    for (auto it = g.neighbors((uint32_t)v); !it.done(); ++it) nb.push_back(*it);
  }
  out.insert(out.end(), nb.begin(), nb.end());
  write_blob(path, out.data(), out.size() * sizeof(uint32_t));
}

double now_s() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

int cmd_rmat(int argc, char **argv) {
  if (argc < 9) return 1;
  uint32_t n = (uint32_t)strtoul(argv[2], 0, 10);
  uint32_t seed = (uint32_t)strtoul(argv[3], 0, 10);
  double a = atof(argv[4]), b = atof(argv[5]), c = atof(argv[6]);
  uint64_t count = strtoull(argv[7], 0, 10);
  rMat<uint32_t> r(n, seed, a, b, c);
  std::vector<uint32_t> out(2 * count);
  for (uint64_t i = 0; i < count; i++) {
    auto e = r((uint32_t)i);
    out[2 * i] = e.first;
    out[2 * i + 1] = e.second;
  }
  write_blob(argv[8], out.data(), out.size() * sizeof(uint32_t));
  return 0;
}

int cmd_run(int argc, char **argv) {
  if (argc < 6) return 1;
  EdgeFile base;
  if (!read_edges(argv[2], base)) { fprintf(stderr, "bad graph file\n"); return 2; }
  std::string prefix = argv[3];
  uint32_t src = (uint32_t)strtoul(argv[4], 0, 10);
  long iters = strtol(argv[5], 0, 10);
  uint64_t n = base.n;
  Graph g((uint32_t)n);
  insert(g, base);
  dump_adjacency(g, n, prefix + ".adj0");
  dump_btree(g, n, prefix + ".btree0");
  for (int k = 6; k < argc; k++) {
    EdgeFile b;
    if (!read_edges(argv[k], b)) { fprintf(stderr, "bad batch file\n"); return 2; }
    insert(g, b);
    dump_adjacency(g, n, prefix + ".adj" + std::to_string(k - 5));
    dump_btree(g, n, prefix + ".btree" + std::to_string(k - 5));
  }
  int32_t *parents = BFS_with_edge_map(g, src);
  write_blob(prefix + ".bfs", parents, n * sizeof(int32_t));
  free(parents);
  double *pr = PR_S<double>(g, iters);
  write_blob(prefix + ".pr", pr, n * sizeof(double));
  free(pr);
  uint32_t *cc = CC(g);
  write_blob(prefix + ".cc", cc, n * sizeof(uint32_t));
  free(cc);
  double *bc = BC(g, src);
  write_blob(prefix + ".bc", bc, n * sizeof(double));
  free(bc);
  uint64_t tc = TC(g);
  write_blob(prefix + ".tc", &tc, sizeof(tc));
  int32_t *sp = SSSP_BF(g, src);
  write_blob(prefix + ".sssp", sp, n * sizeof(int32_t));
  free(sp);
  return 0;
}

int cmd_bench(int argc, char **argv) {
  if (argc < 6) return 1;
  EdgeFile base;
  if (!read_edges(argv[2], base)) { fprintf(stderr, "bad graph file\n"); return 2; }
  uint64_t bs = strtoull(argv[3], 0, 10);
  int batches = atoi(argv[4]);
  uint64_t seed0 = strtoull(argv[5], 0, 10);
  uint64_t n = base.n;
  Graph g((uint32_t)n);
  double t0 = now_s();
  insert(g, base);
  double t_build = now_s() - t0;
  double t_ins = 0;
  uint64_t requested = 0;
  random_aspen r;
  for (int t = 0; t < batches; t++) {
    rMat<uint32_t> gen((uint32_t)n, (uint32_t)r.ith_rand(seed0 + t), 0.5, 0.1, 0.1);
    std::vector<uint64_t> keys;
    keys.reserve(bs);
    for (uint64_t i = 0; i < bs / 2; i++) {
      auto e = gen((uint32_t)i);
      if (e.first == e.second) continue;
      keys.push_back(((uint64_t)e.first << 32) | e.second);
      keys.push_back(((uint64_t)e.second << 32) | e.first);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    EdgeFile b;
    b.n = n;
    for (uint64_t k : keys) { b.src.push_back(k >> 32); b.dst.push_back((uint32_t)k); }
    double t1 = now_s();
    insert(g, b);  // timed region as in graph_bm.cc:359-362 (sort excluded)
    t_ins += now_s() - t1;
    requested += keys.size();
  }
  printf("{\"threads\": %d, \"static_edges\": %zu, \"static_s\": %.6f, "
         "\"batch_size\": %llu, \"batches\": %d, \"inserted\": %llu, "
         "\"insert_s\": %.6f, \"edges_per_s\": %.6e}\n",
         getWorkers(), base.src.size(), t_build, (unsigned long long)bs, batches,
         (unsigned long long)requested, t_ins, requested / t_ins);
  return 0;
}

// The benchmark's mixed step on the reference (src/graph_bm.cc:354-378):
// static build, then per batch add_edge_batch of the insert file (sorted,
// unique; the sort is outside the timed region, as in graph_bm.cc) followed
// by a parallel_for of remove_edge over the delete file.  The delete files
// hold existing edges whose endpoints appear once across all batches, so the
// overlapping-memcpy defect of remove_edge (graph.h:700,735) never meets an
// edge it touched before.
int cmd_mixed(int argc, char **argv) {
  if (argc < 5 || (argc - 3) % 2) return 1;
  EdgeFile base;
  if (!read_edges(argv[2], base)) { fprintf(stderr, "bad graph file\n"); return 2; }
  uint64_t n = base.n;
  Graph g((uint32_t)n);
  double t0 = now_s();
  insert(g, base);
  double t_build = now_s() - t0;
  double t_ins = 0, t_del = 0;
  uint64_t n_ins = 0, n_del = 0;
  int batches = 0;
  for (int k = 3; k + 1 < argc; k += 2) {
    EdgeFile ins, del;
    if (!read_edges(argv[k], ins) || !read_edges(argv[k + 1], del)) { fprintf(stderr, "bad batch file\n"); return 2; }
    double t1 = now_s();
    insert(g, ins);
    double t2 = now_s();
    const uint32_t nd = (uint32_t)del.src.size();
    parallel_for(uint32_t i = 0; i < nd; i++) { g.remove_edge(del.src[i], del.dst[i]); }
    double t3 = now_s();
    t_ins += t2 - t1;
    t_del += t3 - t2;
    n_ins += ins.src.size();
    n_del += nd;
    batches++;
  }
  printf("{\"threads\": %d, \"static_edges\": %zu, \"static_s\": %.6f, \"batches\": %d, "
         "\"inserted\": %llu, \"deleted\": %llu, \"insert_s\": %.6f, \"delete_s\": %.6f, "
         "\"edges_per_s\": %.6e}\n",
         getWorkers(), base.src.size(), t_build, batches, (unsigned long long)n_ins, (unsigned long long)n_del,
         t_ins, t_del, (n_ins + n_del) / (t_ins + t_del));
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_harness rmat|run|bench|mixed ...\n");
    return 1;
  }
  std::string cmd = argv[1];
  int rc = 1;
  if (cmd == "rmat") rc = cmd_rmat(argc, argv);
  else if (cmd == "run") rc = cmd_run(argc, argv);
  else if (cmd == "bench") rc = cmd_bench(argc, argv);
  else if (cmd == "mixed") rc = cmd_mixed(argc, argv);
  if (rc == 1) fprintf(stderr, "bad arguments\n");
  return rc;
}
