// graph_bm: the reference's benchmark driver (src/graph_bm.cc:231-391) on
// the GPU container, written against the graphstore shim.
//
//   graph_bm -s SCALE -m EDGES [-a A -b B -c C] [-src V] [-rounds R]
//            [-maxiters I] [-TC 1] [-SSSP 1] [-weighted 1] [-batches 10,100,...] [-trials T]
//            [-o PREFIX]
//
// The static graph is rMat<uint32_t>(2^SCALE, hash64(100), a, b, c) edges
// 0..EDGES-1 with r = hash32/2^32, self-loops dropped, both directions,
// generated on the device.  Output lines follow the reference's format
// ("AVG\ttest=...", "batch_size = ..., average insert: ...") so its log
// scripts read them.  With -o, BFS parents, PR, CC, BC and TC of the static
// graph are written to PREFIX.{bfs,pr,cc,bc,tc} for checking.
#include <sys/time.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "algorithms.hpp"

using namespace graphstore;

namespace {

double now() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}

uint64_t hash64(uint64_t u) {  // include/rmat_util.h:29-40
  uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
  v ^= v >> 21;
  v ^= v << 37;
  v ^= v >> 4;
  v *= 4768777513237032717ull;
  v ^= v << 20;
  v ^= v >> 41;
  v ^= v << 5;
  return v;
}

struct Args {
  std::map<std::string, std::string> kv;
  std::string get(const std::string &k, const std::string &d) const {
    auto it = kv.find(k);
    return it == kv.end() ? d : it->second;
  }
  long num(const std::string &k, long d) const { return std::stol(get(k, std::to_string(d))); }
  double real(const std::string &k, double d) const { return std::stod(get(k, std::to_string(d))); }
};

template <typename T>
void dump(const std::string &path, const T *p, size_t n) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) {
    perror(path.c_str());
    exit(2);
  }
  fwrite(p, sizeof(T), n, f);
  fclose(f);
}

}  // namespace

// Edge length for the weighted run: symmetric, in [1, 1000].
static uint32_t sym_weight(uint32_t u, uint32_t v) {
  uint64_t a = u < v ? u : v, b = u < v ? v : u;
  uint64_t h = (a * 2654435761ull + b * 40503ull) >> 7;
  return static_cast<uint32_t>(h % 1000 + 1);
}

int main(int argc, char **argv) {
  Args P;
  for (int i = 1; i + 1 < argc; i += 2) P.kv[argv[i]] = argv[i + 1];
  const uint32_t scale = static_cast<uint32_t>(P.num("-s", 16));
  const uint64_t m0 = static_cast<uint64_t>(P.num("-m", 16L << scale));
  const double a = P.real("-a", 0.5), b = P.real("-b", 0.1), c = P.real("-c", 0.1);
  const uint32_t src = static_cast<uint32_t>(P.num("-src", 9));
  const long rounds = P.num("-rounds", 4);
  const long maxiters = P.num("-maxiters", 10);
  const bool with_tc = P.num("-TC", 0) != 0;
  const bool with_sssp = P.num("-SSSP", 0) != 0;  // the test "SSSP_BF" of src/graph_bm.cc:57-84
  const long trials = P.num("-trials", 5);
  const std::string out = P.get("-o", "");
  const uint32_t n = 1u << scale;

  const bool weighted = P.num("-weighted", 0) != 0;  // the WEIGHTED build (src/graph_bm_weighted.cc)
  Graph graph(n, weighted);
  {
    DeviceArray<uint32_t> s(2 * m0), d(2 * m0);
    uint64_t w = 0;
    tg_check(tg_rmat_dev(n, static_cast<uint32_t>(hash64(100)), a, b, c, 1, 0, m0, 3, s.data(), d.data(), &w));
    tg_check(tg_device_sync());
    double t0 = now();
    if (!weighted) {
      tg_check(tg_insert_batch_dev(graph.handle(), s.data(), d.data(), w));
    } else {
      // symmetric lengths in [1, 1000]: w(u, v) = w(v, u)
      std::vector<uint32_t> hs(w), hd(w), hw(w);
      s.download(hs.data(), w);
      d.download(hd.data(), w);
      for (uint64_t i = 0; i < w; i++) hw[i] = sym_weight(hs[i], hd[i]);
      t0 = now();
      std::vector<uint32_t> perm;
      graph.add_edge_batch(hs.data(), hd.data(), hw.data(), static_cast<uint32_t>(w), perm);
    }
    tg_check(tg_device_sync());
    double t = now() - t0;
    printf("Construction finished. Nodes: %u Edges: %llu Size: %.4f GB\n", graph.get_num_vertices(),
           static_cast<unsigned long long>(graph.num_edges64()), graph.get_size() / 1073741824.0);
    printf("Inserting edges: %.6f\nThroughput: %.6e\n", t, graph.num_edges64() / t);
  }

  std::vector<std::string> tests = {"BFS", "PR", "CC", "BC"};
  if (with_tc) tests.push_back("TC");
  if (with_sssp) tests.push_back("SSSP_BF");
  for (const std::string &test : tests) {
    double total = 0;
    for (long i = 0; i < rounds; i++) {
      double t0 = now();
      if (test == "BFS") {
        int32_t *r = BFS_with_edge_map(graph, src);
        total += now() - t0;
        if (!out.empty() && i == 0) dump(out + ".bfs", r, n);
        free(r);
      } else if (test == "PR") {
        double *r = PR_S(graph, maxiters);
        total += now() - t0;
        if (!out.empty() && i == 0) dump(out + ".pr", r, n);
        free(r);
      } else if (test == "CC") {
        uint32_t *r = CC(graph);
        total += now() - t0;
        if (!out.empty() && i == 0) dump(out + ".cc", r, n);
        free(r);
      } else if (test == "BC") {
        double *r = BC(graph, src);
        total += now() - t0;
        if (!out.empty() && i == 0) dump(out + ".bc", r, n);
        free(r);
      } else if (test == "SSSP_BF") {
        int32_t *r = SSSP_BF(graph, src);
        total += now() - t0;
        if (!out.empty() && i == 0) dump(out + ".sssp", r, n);
        free(r);
      } else {
        uint64_t r = TC(graph);
        total += now() - t0;
        if (!out.empty() && i == 0) dump(out + ".tc", &r, 1);
      }
      printf("\ttest=%s\ttime=%f\titeration=%ld\n", test.c_str(), now() - t0, i);
    }
    printf("AVG\ttest=%s\ttime=%f\tgraph=rmat%u\n", test.c_str(), total / rounds, scale);
  }

  if (weighted) return 0;  // the update batches below carry no weights
  // update batches (src/graph_bm.cc:316-390): directed rMAT edges over half
  // the id space, insert then delete the same edges
  std::vector<uint64_t> sizes;
  {
    std::stringstream ss(P.get("-batches", "10,100,1000,10000,100000,1000000"));
    std::string tok;
    while (std::getline(ss, tok, ',')) sizes.push_back(std::stoull(tok));
  }
  const uint32_t nn = n / 2;
  for (uint64_t bs : sizes) {
    double t_ins = 0, t_del = 0;
    DeviceArray<uint32_t> s(bs), d(bs);
    std::vector<uint8_t> present(bs);
    for (long ts = 0; ts < trials; ts++) {
      uint64_t w = 0;
      tg_check(tg_rmat_dev(nn, static_cast<uint32_t>(hash64(100 + ts)), a, b, c, 1, 0, bs, 0, s.data(), d.data(), &w));
      tg_check(tg_device_sync());
      double t0 = now();
      tg_check(tg_insert_batch_dev(graph.handle(), s.data(), d.data(), bs));
      tg_check(tg_device_sync());
      t_ins += now() - t0;
      std::unique_ptr<uint32_t[]> hs = s.to_host(), hd = d.to_host();
      tg_check(tg_is_edge_batch(graph.handle(), hs.get(), hd.get(), bs, present.data()));
      for (uint64_t i = 0; i < bs; i++)
        if (!present[i]) {
          printf("edge (%u, %u) not found, should be\n", hs[i], hd[i]);
          return 1;
        }
      t0 = now();
      tg_check(tg_delete_batch_dev(graph.handle(), s.data(), d.data(), bs));
      tg_check(tg_device_sync());
      t_del += now() - t0;
    }
    double ti = t_ins / trials, td = t_del / trials;
    printf("batch_size = %llu, average insert: %f, throughput %e\n", static_cast<unsigned long long>(bs), ti, bs / ti);
    printf("batch_size = %llu, average delete: %f, throughput %e\n", static_cast<unsigned long long>(bs), td, bs / td);
  }
  return 0;
}
