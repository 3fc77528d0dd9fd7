/*
 * CPU oracle for the Terrace hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's algorithms, used by the
 * tests, by the smoke check and by bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (terrace_amd/, include/terrace_hip.h)
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  The pins that
 * check each function (golden vectors from the reference build, closed
 * forms, brute force) live in tests/test_oracle.py.
 */
#ifndef TERRACE_ORACLE_H
#define TERRACE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generator: include/rmat_util.h:29-115 ---------------------------- */
uint64_t orc_hash64(uint64_t u);
uint32_t orc_hash32(uint32_t a);
/* random_aspen(state).ith_rand(i) = hash64(i + state) (rmat_util.h:57-70) */
uint64_t orc_aspen_ith_rand(uint64_t state, uint64_t i);
/* rMat<uint32_t>(n, seed, a, b, c)(i) for i in [start, start+count).
 * nominal = 0: r = hash32/INT32_MAX exactly as rmat_util.h:95-96 (G_ref);
 * nominal = 1: r = hash32/2^32 (G_nom). Returns 0, or -1 if n is not a
 * power of two or a+b+c > 1 (rmat_util.h:90-91). */
int orc_rmat_edges(uint32_t n, uint32_t seed, double a, double b, double c,
                   int nominal, uint64_t start, uint64_t count,
                   uint32_t *src, uint32_t *dst);

/* ---- edge sets ------------------------------------------------------- */
/* Sort u64 keys ascending (LSD radix, 16-bit digits), in place. */
void orc_sort_u64(uint64_t *keys, uint64_t count);
/* Sort + unique in place; returns the unique count. */
uint64_t orc_sort_unique_u64(uint64_t *keys, uint64_t count);
/* Edge list -> sorted unique directed keys (src<<32|dst); self-loops are
 * dropped; symmetric != 0 adds the reverse of every edge.  out must hold
 * 2*count keys.  Returns the number of keys. */
uint64_t orc_build_keys(const uint32_t *src, const uint32_t *dst,
                        uint64_t count, int symmetric, uint64_t *out);
/* out = a ∪ b (both sorted unique); returns |out|.  out holds na+nb. */
uint64_t orc_set_union(const uint64_t *a, uint64_t na, const uint64_t *b,
                       uint64_t nb, uint64_t *out);
/* out = a \ b (both sorted unique); returns |out|. */
uint64_t orc_set_difference(const uint64_t *a, uint64_t na, const uint64_t *b,
                            uint64_t nb, uint64_t *out);
/* Sorted keys -> CSR (offsets[n+1], nbrs[m]). */
void orc_keys_to_csr(const uint64_t *keys, uint64_t m, uint32_t n,
                     uint64_t *offsets, uint32_t *nbrs);
/* The sorted, de-duplicated CSR of an edge list (both directions when
 * symmetric, self-loops dropped when drop_loops); nbr needs room for every
 * entry before de-duplication.  Returns m. */
uint64_t orc_edges_to_csr(uint32_t n, const uint32_t *src, const uint32_t *dst, uint64_t count, int symmetric,
                          int drop_loops, uint64_t *off, uint32_t *nbr);
/* (CSR ∪ ins) \ del as a new CSR; ins and del sorted unique keys; nbr_out
 * needs room for m + nins.  Returns the new m. */
uint64_t orc_csr_update(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint64_t *ins,
                        uint64_t nins, const uint64_t *del, uint64_t ndel, uint64_t *off_out,
                        uint32_t *nbr_out);

/* ---- tier model: include/graph.h:34-39, 256-372, 379-426, 434-502 ----- */
/* Per-vertex state a reference Graph keeps for tier routing. */
typedef struct {
  uint32_t degree;
  uint8_t btree; /* aux_neighbors != nullptr; never cleared (graph.h) */
} orc_tier_t;
/* Apply one insert batch (sorted unique keys) to the tier model, given the
 * adjacency BEFORE the batch as CSR.  Updates degree with the number of
 * genuinely new edges and sets btree when the run is routed to the B-tree
 * (aux != null || degree + run > MEDIUM_DEGREE, graph.h:276-277,344-345). */
void orc_tier_insert(orc_tier_t *tiers, uint32_t n, const uint64_t *old_off,
                     const uint32_t *old_nbrs, const uint64_t *keys,
                     uint64_t count);

/* ---- algorithms over a symmetric CSR --------------------------------- */
/* BFS levels (include/BFS.h:41-84 VERIFY depths): UINT32_MAX = unreached. */
void orc_bfs_levels(uint32_t n, const uint64_t *off, const uint32_t *nbr,
                    uint32_t src, uint32_t *levels);
/* Check a parent array against levels: parents[src]==src, every reached v
 * has a neighbour parent with level-1, unreached have -1.  Returns the
 * number of violations (0 = valid). */
uint64_t orc_bfs_check_parents(uint32_t n, const uint64_t *off,
                               const uint32_t *nbr, uint32_t src,
                               const int32_t *parents, const uint32_t *levels);
/* PR_S<double> (include/Pagerank.h:106-184): undamped, sequential sums in
 * sorted neighbour order. */
void orc_pagerank(uint32_t n, const uint64_t *off, const uint32_t *nbr,
                  uint32_t iters, double *out);
/* CC (include/Components.h:92-120): minimum vertex id per component. */
void orc_cc(uint32_t n, const uint64_t *off, const uint32_t *nbr,
            uint32_t *labels);
/* BC (include/BC.h:99-158): Brandes in Ligra's inverted form; unreachable
 * vertices come out as (0 - inf)/inf = NaN. */
void orc_bc(uint32_t n, const uint64_t *off, const uint32_t *nbr,
            uint32_t src, double *out);
/* TC (include/TC.h:51-87 + graph.h:1085 count_common): number of
 * triangles w < d < s counted once. */
uint64_t orc_tc(uint32_t n, const uint64_t *off, const uint32_t *nbr);
/* count_common(s, d) = |N(s) ∩ N(d) ∩ [0, min(s,d))| (TC.h:31-47). */
uint32_t orc_count_common(const uint64_t *off, const uint32_t *nbr,
                          uint32_t s, uint32_t d);

/* Triangles through v (the local triangle count query of the serving
 * workload; no counterpart in the reference). */
uint64_t orc_ltc(const uint64_t *off, const uint32_t *nbr, uint32_t v);

/* SSSP_BF with unit edge lengths (include/BellmanFordUnweighted.h:60-130):
 * dist[v] = hop distance from src, INT_MAX/2 when unreachable. */
void orc_sssp_bf(uint32_t n, const uint64_t *off, const uint32_t *nbr, uint32_t src, int32_t *dist);

/* SSSP_BF of the weighted build (include/BellmanFord.h:67-143): w[e] is the
 * length of CSR entry e (edge off[s] <= e < off[s+1] from s); sums saturate
 * at INT_MAX/2 (unreachable). */
void orc_sssp_bf_weighted(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint32_t *w, uint32_t src,
                          int32_t *dist);

#ifdef __cplusplus
}
#endif
#endif
