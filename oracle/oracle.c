/*
 * CPU oracle for the Terrace hot path — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the reference's generator, container semantics and
 * algorithms.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker.  The product never
 * links it.  Each function names the reference file:line it restates; the
 * checks that pin each one are listed in tests/test_oracle.py.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INPLACE 13u          /* include/graph.h:34-39 (unweighted) */
#define MEDIUM_DEGREE 1024u  /* include/graph.h:39 */

/* ---------------------------------------------------------------------- */
/* Generator: include/rmat_util.h:29-115                                  */
/* ---------------------------------------------------------------------- */

uint64_t orc_hash64(uint64_t u) { /* rmat_util.h:29-40 */
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

uint32_t orc_hash32(uint32_t a) { /* rmat_util.h:44-52 */
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

uint64_t orc_aspen_ith_rand(uint64_t state, uint64_t i) { /* :64 */
  return orc_hash64(i + state);
}

int orc_rmat_edges(uint32_t n, uint32_t seed, double a, double b, double c,
                   int nominal, uint64_t start, uint64_t count,
                   uint32_t *src, uint32_t *dst) {
  /* rMat ctor (rmat_util.h:85-92): ab = a+b, abc = a+b+c, h = hash32(seed) */
  double ab = a + b, abc = a + b + c;
  if (abc > 1.0) return -1;
  if (n == 0 || (n & (n - 1)) != 0) return -1;
  uint32_t h = orc_hash32(seed);
  uint32_t levels = 0;
  while ((1u << levels) < n) levels++;
  const double denom = nominal ? 4294967296.0 : 2147483647.0; /* :95-96 */
  /* every edge is a closed form of its index: the loop splits freely */
#pragma omp parallel for schedule(static)
  for (uint64_t e = 0; e < count; e++) {
    uint32_t i = (uint32_t)(start + e);
    /* operator() (rmat_util.h:110-114): 32-bit wrapping products */
    uint32_t rs = orc_hash32((uint32_t)(2u * i) * h);
    uint32_t st = orc_hash32((uint32_t)(2u * i + 1u) * h);
    /* rMatRec (rmat_util.h:98-108) unrolled: recursion depth k uses
     * randStart + k*randStride and splits the block of size n>>k. */
    uint32_t x = 0, y = 0;
    for (uint32_t k = 0; k < levels; k++) {
      uint32_t half = (n >> k) >> 1;
      double r = (double)orc_hash32(rs + k * st) / denom;
      if (r < a) {
      } else if (r < ab) {
        y += half;
      } else if (r < abc) {
        x += half;
      } else {
        x += half;
        y += half;
      }
    }
    src[e] = x;
    dst[e] = y;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Edge sets                                                              */
/* ---------------------------------------------------------------------- */

/* LSD radix sort, 11-bit digits (2048 buckets per chunk stay in cache
 * while scattering).  Each pass is stable: the input is cut
 * into T contiguous chunks, chunk t counts its digits, the scatter offset of
 * (digit d, chunk t) is the number of keys with a smaller digit plus those
 * with digit d in chunks before t, and each chunk scatters in index order.
 * The result does not depend on T. */
void orc_sort_u64(uint64_t *keys, uint64_t count) {
  if (count < 2) return;
  int T = 1;
#ifdef _OPENMP
  T = omp_get_max_threads();
  if ((uint64_t)T > count / 65536 + 1) T = (int)(count / 65536 + 1);  /* small inputs: fewer chunks */
#endif
  uint64_t *tmp = (uint64_t *)malloc(count * sizeof(uint64_t));
  enum { RB = 11, RD = 1 << RB };
  uint64_t *cnt = (uint64_t *)malloc((size_t)T * RD * sizeof(uint64_t));
  uint64_t *from = keys, *to = tmp;
  uint64_t all_or = 0;
#pragma omp parallel for reduction(| : all_or) schedule(static) num_threads(T)
  for (uint64_t i = 0; i < count; i++) all_or |= keys[i];
  for (int shift = 0; shift < 64; shift += RB) {
    if (((all_or >> shift) & (RD - 1)) == 0) continue;
#pragma omp parallel num_threads(T)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      const uint64_t lo = count * (uint64_t)t / (uint64_t)T, hi = count * (uint64_t)(t + 1) / (uint64_t)T;
      uint64_t *c = cnt + (size_t)t * RD;
      memset(c, 0, RD * sizeof(uint64_t));
      for (uint64_t i = lo; i < hi; i++) c[(from[i] >> shift) & (RD - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        uint64_t s = 0;
        for (int d = 0; d < RD; d++)
          for (int u = 0; u < T; u++) {
            uint64_t x = cnt[(size_t)u * RD + d];
            cnt[(size_t)u * RD + d] = s;
            s += x;
          }
      }
      for (uint64_t i = lo; i < hi; i++) to[c[(from[i] >> shift) & (RD - 1)]++] = from[i];
    }
    uint64_t *x = from;
    from = to;
    to = x;
  }
  if (from != keys) memcpy(keys, from, count * sizeof(uint64_t));
  free(tmp);
  free(cnt);
}

uint64_t orc_sort_unique_u64(uint64_t *keys, uint64_t count) {
  if (count == 0) return 0;
  orc_sort_u64(keys, count);
  uint64_t w = 1;
  for (uint64_t i = 1; i < count; i++)
    if (keys[i] != keys[w - 1]) keys[w++] = keys[i];
  return w;
}

uint64_t orc_build_keys(const uint32_t *src, const uint32_t *dst,
                        uint64_t count, int symmetric, uint64_t *out) {
  uint64_t w = 0;
  for (uint64_t i = 0; i < count; i++) {
    if (src[i] == dst[i]) continue;
    out[w++] = ((uint64_t)src[i] << 32) | dst[i];
    if (symmetric) out[w++] = ((uint64_t)dst[i] << 32) | src[i];
  }
  return orc_sort_unique_u64(out, w);
}

uint64_t orc_set_union(const uint64_t *a, uint64_t na, const uint64_t *b,
                       uint64_t nb, uint64_t *out) {
  uint64_t i = 0, j = 0, w = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) out[w++] = a[i++];
    else if (a[i] > b[j]) out[w++] = b[j++];
    else { out[w++] = a[i++]; j++; }
  }
  while (i < na) out[w++] = a[i++];
  while (j < nb) out[w++] = b[j++];
  return w;
}

uint64_t orc_set_difference(const uint64_t *a, uint64_t na, const uint64_t *b,
                            uint64_t nb, uint64_t *out) {
  uint64_t i = 0, j = 0, w = 0;
  while (i < na) {
    while (j < nb && b[j] < a[i]) j++;
    if (j < nb && b[j] == a[i]) { i++; continue; }
    out[w++] = a[i++];
  }
  return w;
}

void orc_keys_to_csr(const uint64_t *keys, uint64_t m, uint32_t n,
                     uint64_t *offsets, uint32_t *nbrs) {
  memset(offsets, 0, ((uint64_t)n + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; i++) offsets[(keys[i] >> 32) + 1]++;
  for (uint32_t v = 0; v < n; v++) offsets[v + 1] += offsets[v];
  for (uint64_t i = 0; i < m; i++) nbrs[i] = (uint32_t)keys[i];
}

static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}

/* The CSR of the edge set {(src[i], dst[i])} (plus the reverse edges when
 * symmetric, self-loops dropped when drop_loops): rows sorted, duplicates
 * removed.  Equals orc_keys_to_csr of orc_build_keys's key set, without the
 * sort of 64-bit keys: each row is counted, filled and sorted on its own.
 * nbr holds the final edges and needs room for every entry before
 * de-duplication (count, or 2 count when symmetric).  Returns m. */
uint64_t orc_edges_to_csr(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t count, int symmetric,
                          int drop_loops, uint64_t *off, uint32_t *nbr) {
  uint32_t *deg = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; i++) {
    if (drop_loops && src[i] == dst[i]) continue;
    __atomic_fetch_add(&deg[src[i]], 1u, __ATOMIC_RELAXED);
    if (symmetric) __atomic_fetch_add(&deg[dst[i]], 1u, __ATOMIC_RELAXED);
  }
  uint64_t *start = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  start[0] = 0;
  for (uint32_t v = 0; v < n; v++) start[v + 1] = start[v] + deg[v];
  uint64_t *cur = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
  memcpy(cur, start, (size_t)n * sizeof(uint64_t));
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < count; i++) {
    if (drop_loops && src[i] == dst[i]) continue;
    nbr[__atomic_fetch_add(&cur[src[i]], 1ull, __ATOMIC_RELAXED)] = dst[i];
    if (symmetric) nbr[__atomic_fetch_add(&cur[dst[i]], 1ull, __ATOMIC_RELAXED)] = src[i];
  }
  free(cur);
  /* sort and de-duplicate each row in place; deg becomes the unique count */
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t v = 0; v < n; v++) {
    uint32_t *r = nbr + start[v];
    uint32_t k = deg[v], w = 0;
    if (k > 1) qsort(r, k, sizeof(uint32_t), cmp_u32);
    for (uint32_t j = 0; j < k; j++)
      if (w == 0 || r[j] != r[w - 1]) r[w++] = r[j];
    deg[v] = w;
  }
  off[0] = 0;
  for (uint32_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
  /* rows move left only (off[v] <= start[v]) and in order, so a serial pass
   * never overwrites a row not yet moved */
  for (uint32_t v = 0; v < n; v++)
    if (off[v] != start[v]) memmove(nbr + off[v], nbr + start[v], (size_t)deg[v] * sizeof(uint32_t));
  free(start);
  free(deg);
  return off[n];
}

/* (CSR ∪ ins) \ del as a new CSR (ins, del: sorted unique keys).  Rows are
 * merged independently.  nbr_out needs room for m + nins. */
uint64_t orc_csr_update(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint64_t *ins,
                        uint64_t nins, const uint64_t *del, uint64_t ndel, uint64_t *off_out,
                        uint32_t *nbr_out) {
  uint64_t *ib = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  uint64_t *db = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  uint64_t *cnt = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
  /* ib[v], db[v]: first insert / delete key with source >= v */
  uint64_t a = 0, b = 0;
  for (uint32_t v = 0; v <= n; v++) {
    while (a < nins && (ins[a] >> 32) < v) a++;
    while (b < ndel && (del[b] >> 32) < v) b++;
    ib[v] = a;
    db[v] = b;
  }
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; v++) {
      uint64_t i = off[v], ie = off[v + 1], j = ib[v], je = ib[v + 1], d = db[v], de = db[v + 1], w = 0;
      uint32_t *out = pass ? nbr_out + off_out[v] : NULL;
      while (i < ie || j < je) {
        uint32_t x;
        if (j >= je || (i < ie && nbr[i] < (uint32_t)ins[j])) x = nbr[i++];
        else if (i >= ie || (uint32_t)ins[j] < nbr[i]) x = (uint32_t)ins[j++];
        else { x = nbr[i++]; j++; }
        while (d < de && (uint32_t)del[d] < x) d++;
        if (d < de && (uint32_t)del[d] == x) continue;
        if (out) out[w] = x;
        w++;
      }
      if (!pass) cnt[v] = w;
    }
    if (!pass) {
      off_out[0] = 0;
      for (uint32_t v = 0; v < n; v++) off_out[v + 1] = off_out[v] + cnt[v];
    }
  }
  free(ib);
  free(db);
  free(cnt);
  return off_out[n];
}

/* ---------------------------------------------------------------------- */
/* Tier model: include/graph.h:256-372 (add_inplace), :434-502            */
/* ---------------------------------------------------------------------- */

void orc_tier_insert(orc_tier_t *tiers, uint32_t n, const uint64_t *old_off,
                     const uint32_t *old_nbrs, const uint64_t *keys,
                     uint64_t count) {
  (void)n;
  uint64_t i = 0;
  while (i < count) {
    uint32_t s = (uint32_t)(keys[i] >> 32);
    uint64_t j = i;
    while (j < count && (uint32_t)(keys[j] >> 32) == s) j++;
    uint64_t size = j - i; /* run length: graph.h:262 */
    uint32_t D = tiers[s].degree;
    const uint32_t *adj = old_nbrs + old_off[s];
    uint64_t deg_old = old_off[s + 1] - old_off[s];
    /* k = |inline ∪ run| where inline = the min(D,13) smallest (:290-330) */
    uint64_t lim = D < INPLACE ? D : INPLACE;
    uint64_t k = 0, p = 0, q = i;
    while (p < lim && q < j) {
      uint32_t d = (uint32_t)keys[q];
      if (adj[p] < d) p++;
      else if (adj[p] > d) q++;
      else { p++; q++; }
      k++;
    }
    k += (lim - p) + (j - q);
    /* routing condition (graph.h:276-277 and :344-345) */
    int overflow = (D == 0) ? (size > INPLACE) : (k > INPLACE);
    if (D == 0) {
      if (tiers[s].btree || D + size > MEDIUM_DEGREE) tiers[s].btree = 1;
    } else if (overflow) {
      if (tiers[s].btree || D + size > MEDIUM_DEGREE) tiers[s].btree = 1;
    }
    /* new degree = |adj ∪ run| (successful inserts only, :363-366, :417) */
    uint64_t newc = 0;
    p = 0;
    for (q = i; q < j; q++) {
      uint32_t d = (uint32_t)keys[q];
      while (p < deg_old && adj[p] < d) p++;
      if (!(p < deg_old && adj[p] == d)) newc++;
    }
    tiers[s].degree = (uint32_t)(D + newc);
    i = j;
  }
}

/* ---------------------------------------------------------------------- */
/* Algorithms                                                             */
/* ---------------------------------------------------------------------- */

void orc_bfs_levels(uint32_t n, const uint64_t *off, const uint32_t *nbr,
                    uint32_t src, uint32_t *levels) {
  /* BFS.h:41-84: level of v = hops from src; UINT32_MAX if unreached. */
  for (uint32_t v = 0; v < n; v++) levels[v] = UINT32_MAX;
  if (src >= n) return;
  uint32_t *q = (uint32_t *)malloc((uint64_t)n * sizeof(uint32_t));
  uint64_t head = 0, tail = 0;
  q[tail++] = src;
  levels[src] = 0;
  while (head < tail) {
    uint32_t u = q[head++];
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t v = nbr[e];
      if (levels[v] == UINT32_MAX) {
        levels[v] = levels[u] + 1;
        q[tail++] = v;
      }
    }
  }
  free(q);
}

static int csr_has(const uint64_t *off, const uint32_t *nbr, uint32_t s,
                   uint32_t d) {
  uint64_t lo = off[s], hi = off[s + 1];
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (nbr[mid] < d) lo = mid + 1;
    else hi = mid;
  }
  return lo < off[s + 1] && nbr[lo] == d;
}

uint64_t orc_bfs_check_parents(uint32_t n, const uint64_t *off,
                               const uint32_t *nbr, uint32_t src,
                               const int32_t *parents,
                               const uint32_t *levels) {
  uint64_t bad = 0;
  for (uint32_t v = 0; v < n; v++) {
    if (levels[v] == UINT32_MAX) {
      if (parents[v] != -1) bad++;
      continue;
    }
    if (v == src) {
      if (parents[v] != (int32_t)src) bad++;
      continue;
    }
    int32_t p = parents[v];
    if (p < 0 || (uint32_t)p >= n) { bad++; continue; }
    if (levels[p] == UINT32_MAX || levels[p] + 1 != levels[v]) { bad++; continue; }
    if (!csr_has(off, nbr, (uint32_t)p, v)) bad++;
  }
  return bad;
}

void orc_pagerank(uint32_t n, const uint64_t *off, const uint32_t *nbr,
                  uint32_t iters, double *out) {
  /* Pagerank.h:106-184: p = 1/n; per iteration p[i] /= deg(i) (PR_Vertex,
   * :61-69), p_next = 0 (:75-84), p_next[i] += p[v] over the sorted
   * neighbours of i (dense pull, :39, Map.cpp:39-48), swap. */
  double *p = out;
  double *q = (double *)malloc((uint64_t)n * sizeof(double));
  double one_over_n = 1 / (double)n;
  for (uint32_t i = 0; i < n; i++) p[i] = one_over_n;
  for (uint32_t it = 0; it < iters; it++) {
    for (uint32_t i = 0; i < n; i++) p[i] = p[i] / (double)(uint32_t)(off[i + 1] - off[i]);
    for (uint32_t i = 0; i < n; i++) {
      double acc = 0.0;
      for (uint64_t e = off[i]; e < off[i + 1]; e++) acc += p[nbr[e]];
      q[i] = acc;
    }
    double *t = p;
    p = q;
    q = t;
  }
  if (p != out) {
    memcpy(out, p, (uint64_t)n * sizeof(double));
    free(p);
  } else {
    free(q);
  }
}

static uint32_t uf_find(uint32_t *par, uint32_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

void orc_cc(uint32_t n, const uint64_t *off, const uint32_t *nbr,
            uint32_t *labels) {
  /* Components.h:92-120 converges to the minimum id of each component;
   * restated by union-find with union-by-minimum. */
  for (uint32_t v = 0; v < n; v++) labels[v] = v;
  for (uint32_t u = 0; u < n; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t a = uf_find(labels, u), b = uf_find(labels, nbr[e]);
      if (a < b) labels[b] = a;
      else if (b < a) labels[a] = b;
    }
  for (uint32_t v = 0; v < n; v++) labels[v] = uf_find(labels, v);
}

void orc_bc(uint32_t n, const uint64_t *off, const uint32_t *nbr,
            uint32_t src, double *out) {
  /* BC.h:99-158.  Forward: NumPaths over BFS levels (BC_F, :34-51);
   * NumPaths[i] = 1/NumPaths[i] (:126-128); backward from the deepest
   * level: Dependencies[d] += Dependencies[s] for s one level deeper
   * (BC_Back_F, :53-71) then += inverseNumPaths[d] (BC_Back_Vertex_F,
   * :84-94); final (D - 1/σ)/(1/σ) (:140-142). */
  uint32_t *lev = (uint32_t *)malloc((uint64_t)n * sizeof(uint32_t));
  uint32_t *order = (uint32_t *)malloc((uint64_t)n * sizeof(uint32_t));
  double *sigma = (double *)calloc(n, sizeof(double));
  double *dep = out;
  for (uint32_t v = 0; v < n; v++) { lev[v] = UINT32_MAX; dep[v] = 0.0; }
  uint64_t head = 0, tail = 0;
  lev[src] = 0;
  sigma[src] = 1.0;
  order[tail++] = src;
  /* level-synchronous: σ[d] = Σ σ[s] over sorted neighbours s one level up */
  uint64_t level_start = 0;
  while (level_start < tail) {
    uint64_t level_end = tail;
    for (head = level_start; head < level_end; head++) {
      uint32_t u = order[head];
      for (uint64_t e = off[u]; e < off[u + 1]; e++) {
        uint32_t v = nbr[e];
        if (lev[v] == UINT32_MAX) {
          lev[v] = lev[u] + 1;
          order[tail++] = v;
        }
      }
    }
    for (uint64_t t = level_end; t < tail; t++) {
      uint32_t v = order[t];
      double acc = 0.0;
      for (uint64_t e = off[v]; e < off[v + 1]; e++)
        if (lev[nbr[e]] + 1 == lev[v]) acc += sigma[nbr[e]];
      sigma[v] = acc;
    }
    level_start = level_end;
  }
  for (uint32_t v = 0; v < n; v++) sigma[v] = 1 / sigma[v];
  /* backward: process vertices deepest level first */
  for (uint64_t t = tail; t-- > 0;) {
    uint32_t v = order[t];
    double acc = 0.0;
    for (uint64_t e = off[v]; e < off[v + 1]; e++)
      if (lev[nbr[e]] == lev[v] + 1) acc += dep[nbr[e]];
    dep[v] = acc + sigma[v];
  }
  for (uint32_t v = 0; v < n; v++) dep[v] = (dep[v] - sigma[v]) / sigma[v];
  free(lev);
  free(order);
  free(sigma);
}

uint32_t orc_count_common(const uint64_t *off, const uint32_t *nbr,
                          uint32_t s, uint32_t d) {
  /* TC.h:31-47 / graph.h:1085: merge while both sides are below their own
   * vertex id. */
  uint64_t i = off[s], ie = off[s + 1], j = off[d], je = off[d + 1];
  uint32_t c = 0;
  while (i < ie && j < je && nbr[i] < s && nbr[j] < d) {
    if (nbr[i] == nbr[j]) { c++; i++; j++; }
    else if (nbr[i] < nbr[j]) i++;
    else j++;
  }
  return c;
}

uint64_t orc_tc(uint32_t n, const uint64_t *off, const uint32_t *nbr) {
  /* TC.h:73-87: edgeMap over all vertices; countF::update(v, i) counts
   * count_common(v, i) for every edge with v > i. */
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (uint32_t i = 0; i < n; i++)
    for (uint64_t e = off[i]; e < off[i + 1]; e++)
      if (nbr[e] > i) total += orc_count_common(off, nbr, nbr[e], i);
  return total;
}

uint64_t orc_ltc(const uint64_t *off, const uint32_t *nbr, uint32_t v) {
  /* triangles through v: half of Σ_{u ∈ N(v)} |N(u) ∩ N(v) \ {u, v}| */
  uint64_t total = 0;
  for (uint64_t e = off[v]; e < off[v + 1]; e++) {
    uint32_t u = nbr[e];
    if (u == v) continue;
    uint64_t i = off[u], ie = off[u + 1], j = off[v], je = off[v + 1];
    while (i < ie && j < je) {
      if (nbr[i] == nbr[j]) {
        if (nbr[i] != u && nbr[i] != v) total++;
        i++;
        j++;
      } else if (nbr[i] < nbr[j]) {
        i++;
      } else {
        j++;
      }
    }
  }
  return total / 2;
}

void orc_sssp_bf(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t src, int32_t *dist) {
  /* BellmanFordUnweighted.h:60-130: ShortestPathLen = INT_MAX/2, 0 at the
   * source; each round relaxes every out-edge of the frontier (BF_F::update,
   * :31-41: dist[s] + 1 < dist[d] lowers dist[d], and d joins the next
   * frontier once per round via Visited), then Visited is reset for the new
   * frontier (BF_Vertex_F, :55-63).  Round n means a negative cycle, which
   * unit lengths cannot produce. */
  const int32_t inf = INT32_MAX / 2;
  uint8_t *visited = (uint8_t *)calloc(n ? n : 1, 1);
  uint32_t *cur = (uint32_t *)malloc((uint64_t)(n ? n : 1) * sizeof(uint32_t));
  uint32_t *nxt = (uint32_t *)malloc((uint64_t)(n ? n : 1) * sizeof(uint32_t));
  for (uint32_t v = 0; v < n; v++) dist[v] = inf;
  dist[src] = 0;
  uint64_t nc = 1, round = 0;
  cur[0] = src;
  while (nc > 0) {
    if (round == n) {
      for (uint32_t v = 0; v < n; v++) dist[v] = -inf;
      break;
    }
    uint64_t nn = 0;
    for (uint64_t i = 0; i < nc; i++) {
      uint32_t s = cur[i];
      for (uint64_t e = off[s]; e < off[s + 1]; e++) {
        uint32_t d = nbr[e];
        int32_t nd = dist[s] + 1;
        if (dist[d] > nd) {
          dist[d] = nd;
          if (!visited[d]) {
            visited[d] = 1;
            nxt[nn++] = d;
          }
        }
      }
    }
    for (uint64_t i = 0; i < nn; i++) visited[nxt[i]] = 0;
    uint32_t *t = cur;
    cur = nxt;
    nxt = t;
    nc = nn;
    round++;
  }
  free(visited);
  free(cur);
  free(nxt);
}

void orc_sssp_bf_weighted(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint32_t *w, uint32_t src,
                          int32_t *dist) {
  /* BellmanFord.h:67-143: the rounds of orc_sssp_bf with BF_F::update's
   * edge length (:30-41) taken from the edge's weight. */
  const int32_t inf = INT32_MAX / 2;
  uint8_t *visited = (uint8_t *)calloc(n ? n : 1, 1);
  uint32_t *cur = (uint32_t *)malloc((uint64_t)(n ? n : 1) * sizeof(uint32_t));
  uint32_t *nxt = (uint32_t *)malloc((uint64_t)(n ? n : 1) * sizeof(uint32_t));
  for (uint32_t v = 0; v < n; v++) dist[v] = inf;
  dist[src] = 0;
  uint64_t nc = 1, round = 0;
  cur[0] = src;
  while (nc > 0) {
    if (round == n) {
      for (uint32_t v = 0; v < n; v++) dist[v] = -inf;
      break;
    }
    uint64_t nn = 0;
    for (uint64_t i = 0; i < nc; i++) {
      uint32_t s = cur[i];
      for (uint64_t e = off[s]; e < off[s + 1]; e++) {
        uint32_t d = nbr[e];
        int64_t nd64 = (int64_t)dist[s] + w[e];
        int32_t nd = nd64 >= inf ? inf : (int32_t)nd64;
        if (dist[d] > nd) {
          dist[d] = nd;
          if (!visited[d]) {
            visited[d] = 1;
            nxt[nn++] = d;
          }
        }
      }
    }
    for (uint64_t i = 0; i < nn; i++) visited[nxt[i]] = 0;
    uint32_t *t = cur;
    cur = nxt;
    nxt = t;
    nc = nn;
    round++;
  }
  free(visited);
  free(cur);
  free(nxt);
}
