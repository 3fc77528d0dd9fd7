// Degree-tiered dynamic graph on gfx950: batch insert / delete across the
// inline, PMA and B-tree tiers, plus the read-side queries.
//
// Reference semantics restated here (include/graph.h):
//   * add_inplace (graph.h:256-372): the 13 smallest neighbours stay inline,
//     the rest of a source run is "kicked" to the secondary tier; the run
//     goes to the B-tree when the vertex already has one or degree + run
//     length > 1024 (graph.h:276-277,344-345), otherwise to the PMA.
//   * add_btree (graph.h:379-426): a vertex routed to the B-tree for the
//     first time moves its PMA neighbours into its new tree.
//   * remove_edge (graph.h:675-791): deleting an inline neighbour refills
//     inline from the smallest secondary neighbour; B-trees are never
//     demoted.
// The batch is processed as whole-array kernels instead of per-edge locks:
// sort + unique, existence filter, one thread per source run for the
// inline merge, then one pass per secondary tier.
#include "graph.hpp"
#include "select.hpp"

#include <algorithm>
#include <cmath>

namespace tg {

namespace {

constexpr uint64_t kInvalidKey = ~0ull;
constexpr uint8_t kRouteNone = 0, kRoutePma = 1, kRouteTree = 2;

// ------------------------------------------------------------ device search
// The searches below are chains of dependent HBM reads, so each step reads
// 15 or 16 fences (or index keys) at once and narrows the range 16-fold:
// one round trip per 16x instead of one per 2x.

// Last leaf in [lo, hi) whose fence <= key, given fence[lo] <= key (or
// lo == 0) and fence[hi] > key (or hi == the leaf count).  Fences never
// decrease.
__device__ inline uint64_t fence_narrow(const GView &g, uint64_t key, uint64_t lo, uint64_t hi) {
  while (hi - lo > 1) {
    const uint64_t step = (hi - lo + 15) >> 4;
    uint64_t f[15];
#pragma unroll
    for (uint32_t x = 1; x < 16; x++) {
      const uint64_t i = lo + x * step;
      f[x - 1] = i < hi ? g.fence[i] : ~0ull;
    }
    uint64_t nlo = lo;
#pragma unroll
    for (uint32_t x = 1; x < 16; x++)
      if (lo + x * step < hi && f[x - 1] <= key) nlo = lo + x * step;
    hi = nlo + step < hi ? nlo + step : hi;
    lo = nlo;
  }
  return lo;
}

// Leaf holding key: the last leaf whose fence <= key (leaf 0 when none).
__device__ inline uint64_t pma_find_leaf(const GView &g, uint64_t key) { return fence_narrow(g, key, 0, 1ull << g.H); }

// From leaf lo with fence[lo] <= key: fences lo+1 .. lo+16 at once, then
// lo+16, lo+32, .. lo+256 and so on (a vertex's run spans a few leaves, so
// one round trip usually brackets the key), then the narrowing above.
__device__ inline uint64_t fence_gallop(const GView &g, uint64_t key, uint64_t lo) {
  const uint64_t nl = 1ull << g.H;
  uint64_t step = 1, hi = nl;
  while (true) {
    uint64_t f[16];
#pragma unroll
    for (uint32_t x = 1; x <= 16; x++) {
      const uint64_t i = lo + x * step;
      f[x - 1] = i < nl ? g.fence[i] : ~0ull;
    }
    uint32_t below = 0;  // samples inside the array with fence <= key (a prefix)
#pragma unroll
    for (uint32_t x = 1; x <= 16; x++) below += (lo + x * step < nl && f[x - 1] <= key) ? 1u : 0u;
    if (below < 16) {
      const uint64_t h = lo + (below + 1) * step;
      hi = h < nl ? h : nl;
      lo += below * step;
      break;
    }
    lo += 16 * step;
    step <<= 4;
  }
  return fence_narrow(g, key, lo, hi);
}

// The leaf of a key below fence[hi] (hi a leaf index): the 16 fences below
// hi at once, since a key just under a vertex's run (a key leaving its
// inline slots, a new smallest secondary neighbour) usually belongs to the
// leaf before the run's first; else the full search.  Fences do not
// decrease, so the last one <= key inside the window, when the window's
// first is <= key, is the last in the array: the same leaf as the search.
__device__ inline uint64_t fence_back(const GView &g, uint64_t key, uint64_t hi) {
  const uint64_t lo = hi > 16 ? hi - 16 : 0;
  uint64_t f[16];
#pragma unroll
  for (uint32_t x = 0; x < 16; x++) f[x] = lo + x < hi ? g.fence[lo + x] : ~0ull;
  if (hi == 0 || f[0] > key) return lo == 0 ? 0 : pma_find_leaf(g, key);  // leaf 0 when none is <= key
  uint64_t r = lo;
#pragma unroll
  for (uint32_t x = 1; x < 16; x++)
    if (lo + x < hi && f[x] <= key) r = lo + x;
  return r;
}

// The same leaf, searched from the leaf of begin_slot (the vertex's first
// PMA slot) onwards, or just below it.  Any other start (none, a tree id)
// takes the full search, so a stale begin slot costs time, never the
// result.
__device__ inline uint64_t pma_find_leaf_from(const GView &g, uint64_t key, uint64_t begin_slot) {
  const uint64_t lo = begin_slot >> kLeafShift;
  if (lo >= (1ull << g.H)) return pma_find_leaf(g, key);
  if (g.fence[lo] > key) return fence_back(g, key, lo);
  return fence_gallop(g, key, lo);
}

// Leaf of key when it is the begin leaf L0 (fence[L0] <= key < fence[L0 + 1],
// the usual case for a vertex of the PMA tier), else the search above.
__device__ inline uint64_t pma_leaf_of(const GView &g, uint64_t key, uint64_t begin) {
  const uint64_t L0 = begin >> kLeafShift;
  if (L0 + 1 < (1ull << g.H)) {
    const uint64_t f0 = g.fence[L0], f1 = g.fence[L0 + 1];
    if (f0 <= key && key < f1) return L0;
    if (f0 <= key) return fence_gallop(g, key, L0 + 1);
  }
  return pma_find_leaf_from(g, key, begin);
}

// The slot among a leaf's destinations v (16 groups of 4 slots from base)
// whose destination is d and source s, or ~0.
__device__ inline uint64_t leaf_match(const GView &g, const uint4 *v, uint64_t base, uint32_t s, uint32_t d) {
#pragma unroll
  for (uint32_t i = 0; i < kLeafSlots / 4; i++) {
    const uint64_t p = base + 4 * i;
    if (v[i].x == d && g.psrc[p] == s) return p;
    if (v[i].y == d && g.psrc[p + 1] == s) return p + 1;
    if (v[i].z == d && g.psrc[p + 2] == s) return p + 2;
    if (v[i].w == d && g.psrc[p + 3] == s) return p + 3;
  }
  return ~0ull;
}

// Slot of (s,d) in the PMA, or ~0 when absent (begin: the vertex's aux
// word).  The begin leaf's destinations are loaded with its two fences, so
// a key inside the begin leaf costs one round trip (and the source check
// where a destination matches); other keys take the searches above.
// leaf_out (optional): the leaf that holds or would hold the key.
__device__ inline uint64_t pma_slot_of(const GView &g, uint32_t s, uint32_t d, uint64_t begin,
                                       uint32_t *leaf_out = nullptr) {
  const uint64_t key = make_key(s, d);
  const uint64_t L0 = begin >> kLeafShift;
  uint4 v[kLeafSlots / 4];
  uint64_t leaf;
  if (L0 + 1 < (1ull << g.H)) {
    const uint64_t f0 = g.fence[L0], f1 = g.fence[L0 + 1];
    const uint4 *pd = reinterpret_cast<const uint4 *>(g.pdst + (L0 << kLeafShift));
#pragma unroll
    for (uint32_t i = 0; i < kLeafSlots / 4; i++) v[i] = pd[i];
    if (f0 <= key && key < f1) {
      if (leaf_out) *leaf_out = static_cast<uint32_t>(L0);
      return leaf_match(g, v, L0 << kLeafShift, s, d);
    }
    leaf = f0 <= key ? fence_gallop(g, key, L0 + 1) : fence_back(g, key, L0);
  } else {
    leaf = pma_find_leaf_from(g, key, begin);
  }
  if (leaf_out) *leaf_out = static_cast<uint32_t>(leaf);
  const uint64_t base = leaf << kLeafShift;
  const uint4 *pd = reinterpret_cast<const uint4 *>(g.pdst + base);
#pragma unroll
  for (uint32_t i = 0; i < kLeafSlots / 4; i++) v[i] = pd[i];
  return leaf_match(g, v, base, s, d);
}

// Index position j of the leaf block of tree t that may hold d
// (last block whose first key <= d, or 0): 16-way narrowing, 15 index keys
// read at once per step, then the last (at most 16) candidates at once.
__device__ inline uint32_t tree_block_pos(const GView &g, uint32_t t, uint32_t d) {
  const uint64_t off = g.toff[t];
  const uint32_t nb = g.tnblk[t];
  uint32_t lo = 0, hi = nb;  // imin[lo] <= d (or lo == 0), imin[hi] > d (or hi == nb)
  while (hi - lo > 16) {
    const uint32_t step = (hi - lo + 15) >> 4;
    uint32_t m[15];
#pragma unroll
    for (uint32_t x = 1; x < 16; x++) {
      const uint32_t i = lo + x * step;
      m[x - 1] = i < hi ? g.imin[off + i] : ~0u;
    }
    uint32_t nlo = lo;
#pragma unroll
    for (uint32_t x = 1; x < 16; x++)
      if (lo + x * step < hi && m[x - 1] <= d) nlo = lo + x * step;
    hi = nlo + step < hi ? nlo + step : hi;
    lo = nlo;
  }
  uint32_t m[15];
#pragma unroll
  for (uint32_t x = 1; x < 16; x++) m[x - 1] = lo + x < hi ? g.imin[off + lo + x] : ~0u;
  uint32_t j = lo;
#pragma unroll
  for (uint32_t x = 1; x < 16; x++)
    if (lo + x < hi && m[x - 1] <= d) j = lo + x;
  return j;
}

// d among the first cnt keys of a leaf block (kBlk = 256 slots, sorted):
// 16 samples at stride 16, then the 16 keys after the last sample <= d,
// each level loaded at once (slots past cnt are read but never used).
__device__ inline bool block_has(const uint32_t *keys, uint32_t cnt, uint32_t d) {
  static_assert(kBlk == 256, "two 16-way levels cover a block");
  uint32_t smp[16];
#pragma unroll
  for (uint32_t x = 1; x < 16; x++) smp[x] = keys[16 * x];
  uint32_t lo = 0;
#pragma unroll
  for (uint32_t x = 1; x < 16; x++)
    if (16 * x < cnt && smp[x] <= d) lo = 16 * x;
  uint32_t v[16];
#pragma unroll
  for (uint32_t x = 0; x < 16; x++) v[x] = keys[lo + x];
  bool hit = false;
#pragma unroll
  for (uint32_t x = 0; x < 16; x++) hit |= lo + x < cnt && v[x] == d;
  return hit;
}

__device__ inline bool tree_has(const GView &g, uint32_t t, uint32_t d) {
  uint32_t nb = g.tnblk[t];
  if (nb == 0) return false;
  uint32_t j = tree_block_pos(g, t, d);
  uint32_t b = g.iblk[g.toff[t] + j];
  return block_has(g.bkeys + static_cast<uint64_t>(b) * kBlk, g.bcnt[b], d);
}

// Graph::is_edge (graph.h:643-673): inline first, then the secondary tier.
// The 64-byte block is loaded whole (four 16-byte loads issued together).
// leaf_out (optional): the PMA leaf searched for a key of a PMA-tier vertex
// (the leaf that holds or would hold it), else left as it is.
// slot_out (optional): the PMA slot that holds the key when the lookup
// found it there (left unchanged otherwise).
__device__ inline bool edge_exists(const GView &g, uint32_t s, uint32_t d, uint32_t *leaf_out = nullptr,
                                   uint64_t *slot_out = nullptr) {
  const uint4 *vp = reinterpret_cast<const uint4 *>(g.vb + s);
  const uint4 q0 = vp[0], q1 = vp[1], q2 = vp[2], q3 = vp[3];
  // degree, nbr[0..12], aux (low, high)
  const uint32_t w[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                          q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
  const uint32_t D = w[0];
  const uint32_t lim = D < kInline ? D : kInline;
  bool hit = false;
#pragma unroll
  for (uint32_t i = 0; i < kInline; i++) hit |= i < lim && w[1 + i] == d;
  if (hit) return true;
  // inline slots hold the 13 smallest neighbours: past them only if d is larger
  if (D <= kInline || d < w[kInline]) return false;
  const uint64_t aux = static_cast<uint64_t>(w[14]) | (static_cast<uint64_t>(w[15]) << 32);
  if (aux & kAuxTree) return tree_has(g, static_cast<uint32_t>(aux), d);
  const uint64_t sl = pma_slot_of(g, s, d, aux, leaf_out);
  if (slot_out) *slot_out = sl;
  return sl != ~0ull;
}

// Merge path: number of A elements among the first m of merge(A, B)
// (A and B sorted, disjoint).
template <typename TA, typename TB>
__device__ inline uint64_t merge_split(const TA &A, uint64_t na, const TB &B, uint64_t nb, uint64_t m) {
  uint64_t lo = m > nb ? m - nb : 0;
  uint64_t hi = m < na ? m : na;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (A(mid) < B(m - 1 - mid)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

struct KeyArr {
  const uint64_t *p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i]; }
};
struct DstArr {  // dst part of a key array
  const uint64_t *p;
  __device__ uint32_t operator()(uint64_t i) const { return static_cast<uint32_t>(p[i]); }
};
struct U32Arr {
  const uint32_t *p;
  __device__ uint32_t operator()(uint64_t i) const { return p[i]; }
};

template <typename T>
__device__ inline uint64_t lower_bound_dev(const T *a, uint64_t n, T x) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// A count the device computed, bounded by the launch's capacity (null: the
// capacity itself).
__device__ inline uint64_t dev_count(const uint64_t *p, uint64_t cap) {
  if (!p) return cap;
  const uint64_t v = *p;
  return v < cap ? v : cap;
}

#define GRID_STRIDE(i, n)                                                                  \
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < (n); \
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)

// ------------------------------------------------------------ batch prep
// Mixed batches (inserts and deletes in one call) sort tagged keys:
// src << 32 | dst << 1 | op (op 0 insert, 1 delete; ids are below 2^31), so
// the keys of one edge are adjacent, its insert before its delete.
__host__ __device__ inline uint64_t make_tagged(uint32_t s, uint32_t d, uint32_t op) {
  return (static_cast<uint64_t>(s) << 32) | (static_cast<uint64_t>(d) << 1) | op;
}
__host__ __device__ inline uint64_t untag(uint64_t t) {
  return (t & 0xFFFFFFFF00000000ull) | ((t & 0xFFFFFFFFull) >> 1);
}

__global__ void k_pack(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t nv, uint64_t *keys,
                       uint64_t *err) {
  GRID_STRIDE(i, n) {
    uint32_t s = src[i], d = dst[i];
    if (s >= nv || d >= nv) {
      atomicOr(reinterpret_cast<unsigned long long *>(err), 1ull);
      keys[i] = 0;
    } else {
      keys[i] = make_key(s, d);
    }
  }
}

// Batches of at most kSmallSortKeys pairs: pack, validate, sort and unique
// in one workgroup (bitonic sort over LDS, padded with ~0 to a power of
// two); the result equals k_pack, the radix sort and the unique select.
constexpr uint32_t kSmallSortKeys = 4096;

// Exclusive scan of one value per thread over a 1024-thread workgroup (wave
// shuffles, then the 16 wave totals); *total gets the sum.  s_w: 16 words.
__device__ inline uint32_t wg_excl_scan(uint32_t x, uint32_t *s_w, uint32_t *total) {
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t v = x;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) s_w[w] = v;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (uint32_t i = 0; i < 16; i++) {
    const uint32_t c = s_w[i];
    if (i < w) base += c;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return base + v - x;
}

__global__ void __launch_bounds__(1024) k_pack_sort_small(const uint32_t *src, const uint32_t *dst, uint32_t count,
                                                          uint32_t nv, uint64_t *keys, uint64_t *err,
                                                          uint64_t *nunique) {
  __shared__ uint64_t s[kSmallSortKeys];
  __shared__ uint32_t s_w[16];
  const uint32_t t = threadIdx.x;
  if (t < 14) err[t] = 0;  // the batch scalars (the look-back error words [14], [15] stay)
  __syncthreads();
  uint32_t P = 1;
  while (P < count) P <<= 1;
  bool bad = false;
  for (uint32_t i = t; i < P; i += blockDim.x) {
    uint64_t v = ~0ull;
    if (i < count) {
      const uint32_t a = src[i], b = dst[i];
      if (a >= nv || b >= nv) {
        bad = true;
        v = 0;
      } else {
        v = make_key(a, b);
      }
    }
    s[i] = v;
  }
  if (bad) atomicOr(reinterpret_cast<unsigned long long *>(err), 1ull);
  __syncthreads();
  wg_bitonic_sort(s, P);
  // unique, in order: thread t takes keys 4t .. 4t+3
  bool f[4];
  uint32_t c = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; j++) {
    const uint32_t i = 4 * t + j;
    f[j] = i < count && (i == 0 || s[i] != s[i - 1]);
    c += f[j] ? 1u : 0u;
  }
  uint32_t tot;
  uint32_t pos = wg_excl_scan(c, s_w, &tot);
#pragma unroll
  for (uint32_t j = 0; j < 4; j++)
    if (f[j]) keys[pos++] = s[4 * t + j];
  if (t == 0) *nunique = tot;
}

// Mixed batches of at most kSmallSortKeys edges in all: pack, validate,
// sort and split in one workgroup.  The inserts (is, id) and deletes (ds,
// dd) are packed as tagged keys, sorted in LDS and split by the last op of
// each edge (deletes come after inserts, so a batch that inserts and
// deletes an edge leaves it absent, as applying the inserts and then the
// deletes does): edges inserted into I (count sc[1]), edges deleted into D
// (count sc[9]).  Clears the batch scalars.  The lookups, prefixes and runs
// of both halves then run in k_mixed_prep across many workgroups: in this
// one workgroup the thousand lookup chains of a 1e3 batch took 30 us of the
// kernel's 50.
__global__ void __launch_bounds__(1024) k_small_mixed_sort(const uint32_t *is, const uint32_t *id, uint32_t ni,
                                                           const uint32_t *ds, const uint32_t *dd, uint32_t nd,
                                                           uint32_t nv, uint64_t *I, uint64_t *D, uint64_t *sc) {
  __shared__ uint64_t s[kSmallSortKeys];
  __shared__ uint32_t s_w[16];
  const uint32_t t = threadIdx.x;
  const uint32_t count = ni + nd;
  if (t < 14) sc[t] = 0;  // the batch scalars (the look-back error words [14], [15] stay)
  __syncthreads();
  uint32_t P = 1;
  while (P < count) P <<= 1;
  bool bad = false;
  for (uint32_t i = t; i < P; i += blockDim.x) {
    uint64_t v = ~0ull;
    if (i < count) {
      const uint32_t a = i < ni ? is[i] : ds[i - ni], b = i < ni ? id[i] : dd[i - ni];
      if (a >= nv || b >= nv) {
        bad = true;
        v = 0;
      } else {
        v = make_tagged(a, b, i < ni ? 0u : 1u);
      }
    }
    s[i] = v;
  }
  if (bad) atomicOr(reinterpret_cast<unsigned long long *>(sc), 1ull);
  __syncthreads();
  wg_bitonic_sort(s, P);
  // the last key of each edge decides: thread t takes keys 4t .. 4t+3
  bool fi[4], fd[4];
  uint64_t kk[4];
  uint32_t ci = 0, cd = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; j++) {
    const uint32_t i = 4 * t + j;
    const bool last = i < count && (i + 1 == count || (s[i] >> 1) != (s[i + 1] >> 1));
    fi[j] = last && (s[i] & 1) == 0;
    fd[j] = last && (s[i] & 1) != 0;
    kk[j] = i < count ? untag(s[i]) : 0;
    ci += fi[j] ? 1u : 0u;
    cd += fd[j] ? 1u : 0u;
  }
  uint32_t nI, nD;
  uint32_t pi = wg_excl_scan(ci, s_w, &nI);
  uint32_t pd = wg_excl_scan(cd, s_w, &nD);
#pragma unroll
  for (uint32_t j = 0; j < 4; j++) {
    if (fi[j]) I[pi++] = kk[j];
    if (fd[j]) D[pd++] = kk[j];
  }
  if (t == 0) {
    sc[1] = nI;
    sc[9] = nD;
  }
}

// Radix path of mixed batches: tagged keys.
__global__ void k_pack_mixed(const uint32_t *is, const uint32_t *id, uint64_t ni, const uint32_t *ds,
                             const uint32_t *dd, uint64_t nd, uint32_t nv, uint64_t *keys, uint64_t *err) {
  GRID_STRIDE(i, ni + nd) {
    const uint32_t s = i < ni ? is[i] : ds[i - ni], d = i < ni ? id[i] : dd[i - ni];
    if (s >= nv || d >= nv) {
      atomicOr(reinterpret_cast<unsigned long long *>(err), 1ull);
      keys[i] = 0;
    } else {
      keys[i] = make_tagged(s, d, i < ni ? 0u : 1u);
    }
  }
}

// Batches of more than kSmallSortKeys and at most kBucketSortMax pairs: a
// bucket sort.  The compressed key c = src << bits | dst (vertex ids below
// 2^bits) orders like the key; its top 12 bits pick one of 4096 buckets,
// so buckets follow key order.  Pack, validate and count per bucket;
// scatter by bucket (each block scans the counts itself; order inside a
// bucket is arbitrary); one workgroup per bucket sorts it in LDS (bitonic);
// then the unique select as after the radix sort.  A bucket beyond kBucketCap
// keys sets kSortOverflow in the error word and the caller sorts again by
// radix.  The sort kernel clears the counts and cursors for the next call.
constexpr uint32_t kBucketBits = 12, kBuckets = 1u << kBucketBits;
constexpr uint32_t kBucketCap = 2048;
constexpr uint64_t kBucketSortMax = 1u << 18;
constexpr uint64_t kSortOverflow = 256;
constexpr uint32_t kBucketBlock = 256, kBucketItems = 8, kBucketTile = kBucketBlock * kBucketItems;

template <bool kTagged>
__device__ inline uint32_t bucket_of(uint64_t key, uint32_t bits, uint32_t shift) {
  const uint64_t k = kTagged ? untag(key) : key;
  const uint64_t c = (static_cast<uint64_t>(key_src(k)) << bits) | key_dst(k);
  return static_cast<uint32_t>(c >> shift);
}

// Pack, validate and count per bucket.  Items [0, count) come from
// (src, dst); with kTagged, items [count, count + count2) from (src2, dst2)
// with op 1, the first ones with op 0.
template <bool kTagged>
__global__ void __launch_bounds__(kBucketBlock) k_bucket_count(const uint32_t *src, const uint32_t *dst,
                                                               uint32_t count, const uint32_t *src2,
                                                               const uint32_t *dst2, uint32_t count2, uint32_t nv,
                                                               uint32_t bits, uint32_t shift, uint64_t *packed,
                                                               uint32_t *bcnt, uint64_t *err) {
  __shared__ uint32_t s_c[kBuckets];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < kBuckets; i += kBucketBlock) s_c[i] = 0;
  __syncthreads();
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBucketTile;
  const uint64_t total = static_cast<uint64_t>(count) + (kTagged ? count2 : 0u);
  uint32_t ra[kBucketItems], rb[kBucketItems];
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; j++) {
    const uint64_t i = base + j * kBucketBlock + t;
    ra[j] = i < count ? src[i] : (kTagged && i < total ? src2[i - count] : 0);
    rb[j] = i < count ? dst[i] : (kTagged && i < total ? dst2[i - count] : 0);
  }
  bool bad = false;
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; j++) {
    const uint64_t i = base + j * kBucketBlock + t;
    if (i >= total) continue;
    uint64_t key = 0;
    if (ra[j] >= nv || rb[j] >= nv) bad = true;
    else key = kTagged ? make_tagged(ra[j], rb[j], i >= count ? 1u : 0u) : make_key(ra[j], rb[j]);
    packed[i] = key;
    atomicAdd(&s_c[bucket_of<kTagged>(key, bits, shift)], 1u);
  }
  if (bad) atomicOr(reinterpret_cast<unsigned long long *>(err), 1ull);
  __syncthreads();
  for (uint32_t i = t; i < kBuckets; i += kBucketBlock)
    if (s_c[i]) atomicAdd(&bcnt[i], s_c[i]);
}

// Exclusive scan of the 4096 bucket counts into s_off (256 threads, 16
// buckets each); block 0 also stores the starts for the sort kernel.
__device__ inline void bucket_starts(const uint32_t *bcnt, uint32_t *s_off, uint32_t *s_w, uint32_t *boff) {
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x[16], sum = 0;
#pragma unroll
  for (uint32_t q = 0; q < 16; q++) {
    x[q] = bcnt[16 * t + q];
    sum += x[q];
  }
  uint32_t inc = sum;
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t run = inc - sum;
  for (uint32_t v = 0; v < w; v++) run += s_w[v];
#pragma unroll
  for (uint32_t q = 0; q < 16; q++) {
    s_off[16 * t + q] = run;
    if (blockIdx.x == 0) boff[16 * t + q] = run;
    run += x[q];
  }
  __syncthreads();
}

template <bool kTagged>
__global__ void __launch_bounds__(kBucketBlock) k_bucket_scatter(const uint64_t *packed, uint32_t count,
                                                                 uint32_t bits, uint32_t shift,
                                                                 const uint32_t *bcnt, uint32_t *bcur,
                                                                 uint32_t *boff, uint64_t *out) {
  __shared__ uint32_t s_off[kBuckets];
  __shared__ uint32_t s_c[kBuckets];
  __shared__ uint32_t s_w[kBucketBlock / 64];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < kBuckets; i += kBucketBlock) s_c[i] = 0;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBucketTile;
  uint64_t k[kBucketItems];
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; j++) {
    const uint64_t i = base + j * kBucketBlock + t;
    k[j] = i < count ? packed[i] : 0;
  }
  bucket_starts(bcnt, s_off, s_w, boff);  // also orders the s_c clearing before its use
  uint32_t r[kBucketItems];
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; j++) {
    const uint64_t i = base + j * kBucketBlock + t;
    r[j] = i < count ? atomicAdd(&s_c[bucket_of<kTagged>(k[j], bits, shift)], 1u) : 0u;
  }
  __syncthreads();
  // one global reservation per bucket this block holds keys of
  for (uint32_t i = t; i < kBuckets; i += kBucketBlock)
    if (s_c[i]) s_off[i] += atomicAdd(&bcur[i], s_c[i]);
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; j++) {
    const uint64_t i = base + j * kBucketBlock + t;
    if (i >= count) continue;
    const uint32_t pos = s_off[bucket_of<kTagged>(k[j], bits, shift)] + r[j];
    if (pos < count) out[pos] = k[j];  // always, unless the counters were stale
  }
}

// One workgroup per bucket sorts it in LDS (bitonic) and writes it back in
// place; no workgroup waits on another.
__global__ void __launch_bounds__(kBucketBlock) k_bucket_sort(uint64_t *keys, uint32_t count, const uint32_t *boff,
                                                              uint32_t *bcnt, uint32_t *bcur, uint64_t *err) {
  __shared__ uint64_t s[kBucketCap];
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  const uint32_t start = boff[b], size = bcnt[b];
  __syncthreads();  // every thread has read the count before it is cleared
  if (t == 0) {
    bcnt[b] = 0;
    bcur[b] = 0;
  }
  // a bucket past the batch (stale counters) is treated as an overflow too
  const bool fits = size <= kBucketCap && static_cast<uint64_t>(start) + size <= count;
  if (!fits && t == 0) atomicOr(reinterpret_cast<unsigned long long *>(err), kSortOverflow);
  const uint32_t n = fits ? size : 0;
  if (n < 2) return;  // uniform over the workgroup
  uint32_t P = 2;
  while (P < n) P <<= 1;
  for (uint32_t i = t; i < P; i += kBucketBlock) s[i] = i < n ? keys[start + i] : ~0ull;
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = t; i < P; i += kBucketBlock) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const uint64_t x = s[i], y = s[ixj];
          if ((x > y) == ((i & k) == 0)) {
            s[i] = y;
            s[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t i = t; i < n; i += kBucketBlock) keys[start + i] = s[i];
}

// After the existence check of a small batch (at most kSmallSortKeys keys,
// one workgroup).  Insert: nex (flags, 0 past the unique count) becomes its
// exclusive prefix over [0, cap] and the source runs of R are listed, with
// run_beg[nruns] = nR.  Delete: the existing keys of R (ex flags) are
// compacted into E and the source runs of E listed.  Counts into sc[2]
// (insert: runs; delete: existing) and sc[3] (delete: runs); nR is sc[1].
template <bool kInsert>
__global__ void __launch_bounds__(1024) k_small_runs(const uint64_t *R, uint32_t *flags, uint32_t cap, uint64_t *E,
                                                     uint32_t *run_beg, uint64_t *sc) {
  __shared__ uint64_t s_k[kSmallSortKeys];
  __shared__ uint32_t s_w[16];
  const uint32_t t = threadIdx.x;
  const uint32_t nR = static_cast<uint32_t>(sc[1]);
  uint32_t tot;
  uint32_t m = nR;  // keys whose runs are listed (R, or E for deletes)
  if (kInsert) {
    // exclusive prefix of the new-key flags over [0, cap], total at cap
    uint32_t x[4], c = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      const uint32_t i = 4 * t + j;
      x[j] = i < cap ? flags[i] : 0u;
      c += x[j];
    }
    uint32_t pre = wg_excl_scan(c, s_w, &tot);
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      const uint32_t i = 4 * t + j;
      if (i < cap) flags[i] = pre;
      pre += x[j];
    }
    if (t == 0) flags[cap] = tot;
    for (uint32_t i = t; i < nR; i += blockDim.x) s_k[i] = R[i];
  } else {
    // compact the existing keys into E (and LDS)
    bool f[4];
    uint64_t k[4];
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      const uint32_t i = 4 * t + j;
      f[j] = i < nR && flags[i] != 0;
      k[j] = f[j] ? R[i] : 0;
      c += f[j] ? 1u : 0u;
    }
    uint32_t pos = wg_excl_scan(c, s_w, &tot);
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
      if (f[j]) {
        E[pos] = k[j];
        s_k[pos] = k[j];
        pos++;
      }
    }
    m = tot;
    if (t == 0) sc[2] = tot;
  }
  __syncthreads();
  // source runs of s_k[0, m)
  bool f[4];
  uint32_t c = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; j++) {
    const uint32_t i = 4 * t + j;
    f[j] = i < m && (i == 0 || key_src(s_k[i]) != key_src(s_k[i - 1]));
    c += f[j] ? 1u : 0u;
  }
  uint32_t pos = wg_excl_scan(c, s_w, &tot);
#pragma unroll
  for (uint32_t j = 0; j < 4; j++)
    if (f[j]) run_beg[pos++] = 4 * t + j;
  if (t == 0) {
    run_beg[tot] = m;
    sc[kInsert ? 2 : 3] = tot;
  }
}

__global__ void k_scatter_flagged(const uint64_t *in, const uint32_t *flags, const uint32_t *pos, uint64_t n,
                                  uint64_t *out) {
  GRID_STRIDE(i, n) if (flags[i]) out[pos[i]] = in[i];
}

// exists flag per (sorted unique) key
__global__ void k_check_weights(const uint32_t *w, uint64_t n, uint64_t *err) {
  GRID_STRIDE(i, n) if (w[i] > 0x7FFFFFFFu) atomicOr(reinterpret_cast<unsigned long long *>(err), 8ull);
}

__global__ void k_exists(GView g, const uint64_t *keys, uint64_t n, uint32_t *ex) {
  GRID_STRIDE(i, n) ex[i] = edge_exists(g, key_src(keys[i]), key_dst(keys[i])) ? 1u : 0u;
}

__global__ void k_exists2(GView g, const uint64_t *keys, uint64_t n, uint32_t *ex, uint32_t *nex) {
  GRID_STRIDE(i, n) {
    bool e = edge_exists(g, key_src(keys[i]), key_dst(keys[i]));
    ex[i] = e ? 1u : 0u;
    nex[i] = e ? 0u : 1u;
  }
}

// Same, over the first *np keys of a buffer of cap (flags 0 beyond), so
// the unique count need not be read back first.
__global__ void k_exists2_n(GView g, const uint64_t *keys, const uint64_t *np, uint64_t cap, uint32_t *ex,
                            uint32_t *nex) {
  const uint64_t n = *np;
  GRID_STRIDE(i, cap) {
    bool e = i < n && edge_exists(g, key_src(keys[i]), key_dst(keys[i]));
    ex[i] = e ? 1u : 0u;
    nex[i] = (i < n && !e) ? 1u : 0u;
  }
}
__global__ void k_exists_n(GView g, const uint64_t *keys, const uint64_t *np, uint64_t cap, uint32_t *ex) {
  const uint64_t n = *np;
  GRID_STRIDE(i, cap) ex[i] = (i < n && edge_exists(g, key_src(keys[i]), key_dst(keys[i]))) ? 1u : 0u;
}

// Insert preparation of a batch past kSmallSortKeys, after the existence
// flags ex (k_exists2_n, one key per thread: the lookups need the whole
// GPU's memory parallelism), in one single-pass launch over the first *np
// unique keys R (of cap): nidx[i] = new keys before i, for every i <= cap
// (nidx[cap] = their total); the source run starts into run_beg, their
// count into *nruns and run_beg[nruns] = *np.  Two look-back columns (new
// keys, run starts), each walked by its own wave; ranks within the tile
// come from ballots, so both outputs follow index order.
constexpr int kPrepBlock = 256;

// kItems keys per thread.  kLookup: the existence flags are computed here
// (one key per thread keeps the lookups' memory parallelism) and written
// to ex; otherwise ex holds them already.
// The work of block b (in start order) of nblocks.
// leaf (optional, with kLookup): per key the PMA leaf its lookup searched,
// ~0 when the lookup did not reach the PMA.  slot (optional, with
// kLookup): per key the PMA slot that holds it, ~0 when it is not there.
template <int kItems, bool kLookup>
__device__ inline void insert_prep_block(uint32_t b, GView g, const uint64_t *R, const uint64_t *np, uint64_t cap,
                                         uint32_t *ex, uint32_t *nidx, uint32_t *run_beg, uint64_t *nruns,
                                         uint64_t *state, uint32_t *err, uint32_t nblocks, uint32_t epoch,
                                         uint32_t *leaf = nullptr, uint64_t *slot = nullptr) {
  constexpr int kPrepItems = kItems, kPrepTile = kPrepBlock * kItems;
  __shared__ uint32_t wcnt[2][kPrepItems][kPrepBlock / 64];
  __shared__ uint32_t wpre[2][kPrepItems][kPrepBlock / 64];
  __shared__ uint32_t s_run[2];
  __shared__ uint64_t s_off[2];
  const uint64_t n = *np;
  const uint32_t t = threadIdx.x, w = t >> 6;
  const uint64_t lt = lanemask_lt();
  const uint64_t base = static_cast<uint64_t>(b) * kPrepTile;
  bool e[kPrepItems], st[kPrepItems];
#pragma unroll
  for (int i = 0; i < kPrepItems; i++) {
    const uint64_t idx = base + static_cast<uint64_t>(i) * kPrepBlock + t;
    e[i] = false;
    st[i] = false;
    if (idx < n) {
      const uint64_t key = R[idx];
      if (kLookup) {
        uint32_t lf = ~0u;
        uint64_t sl = ~0ull;
        e[i] = edge_exists(g, key_src(key), key_dst(key), leaf ? &lf : nullptr, slot ? &sl : nullptr);
        if (leaf) leaf[idx] = lf;
        if (slot) slot[idx] = sl;
      } else {
        e[i] = ex[idx] != 0;
      }
      st[i] = idx == 0 || key_src(R[idx - 1]) != key_src(key);
    }
  }
  uint32_t rn[kPrepItems], rs[kPrepItems];
#pragma unroll
  for (int i = 0; i < kPrepItems; i++) {
    const uint64_t idx = base + static_cast<uint64_t>(i) * kPrepBlock + t;
    const uint64_t mn = __ballot(idx < n && !e[i]), ms = __ballot(st[i]);
    rn[i] = static_cast<uint32_t>(__popcll(mn & lt));
    rs[i] = static_cast<uint32_t>(__popcll(ms & lt));
    if ((t & 63) == 0) {
      wcnt[0][i][w] = static_cast<uint32_t>(__popcll(mn));
      wcnt[1][i][w] = static_cast<uint32_t>(__popcll(ms));
    }
  }
  __syncthreads();
  if (t < 128) {  // wave 0 walks column 0, wave 1 column 1
    const uint32_t col = t >> 6;
    const bool lead = (t & 63) == 0;
    uint32_t run = 0;
    for (int i = 0; i < kPrepItems; i++)
      for (int v = 0; v < kPrepBlock / 64; v++) {
        if (lead) wpre[col][i][v] = run;
        run += wcnt[col][i][v];
      }
    const uint32_t off = lbk::lookback_wave(state, 2, b, col, err, epoch, run);
    if (lead) {
      s_off[col] = off;
      s_run[col] = run;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kPrepItems; i++) {
    const uint64_t idx = base + static_cast<uint64_t>(i) * kPrepBlock + t;
    if (idx >= cap) continue;
    if (kLookup) ex[idx] = e[i] ? 1u : 0u;
    nidx[idx] = static_cast<uint32_t>(s_off[0] + wpre[0][i][w] + rn[i]);
    if (st[i]) run_beg[s_off[1] + wpre[1][i][w] + rs[i]] = static_cast<uint32_t>(idx);
  }
  if (b == nblocks - 1 && t == 0) {
    const uint64_t runs = s_off[1] + s_run[1];
    nidx[cap] = static_cast<uint32_t>(s_off[0] + s_run[0]);
    *nruns = runs;
    run_beg[runs] = static_cast<uint32_t>(n);
  }
}

template <int kItems, bool kLookup>
__global__ void __launch_bounds__(kPrepBlock) k_insert_prep(GView g, const uint64_t *R, const uint64_t *np,
                                                            uint64_t cap, uint32_t *ex, uint32_t *nidx,
                                                            uint32_t *run_beg, uint64_t *nruns, uint64_t *state,
                                                            uint32_t *misc, uint32_t *err, uint32_t nblocks,
                                                            uint32_t epoch) {
  __shared__ uint32_t s_ticket;
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  insert_prep_block<kItems, kLookup>(b, g, R, np, cap, ex, nidx, run_beg, nruns, state, err, nblocks, epoch);
}

// Delete preparation of a batch of kSmallSortKeys to kBucketSortMax keys in
// one single-pass launch: the existing keys among the first *np of R
// compacted into out in order (count into *nout), one key per thread (the
// lookups keep their memory parallelism).
__device__ inline void delete_prep_block(uint32_t b, GView g, const uint64_t *R, const uint64_t *np, uint64_t *out,
                                         uint64_t *nout, uint64_t *state, uint32_t *err, uint32_t nblocks,
                                         uint32_t epoch) {
  __shared__ uint32_t wcnt[kPrepBlock / 64];
  __shared__ uint64_t s_off;
  const uint64_t n = *np;
  const uint32_t t = threadIdx.x, w = t >> 6;
  const uint64_t idx = static_cast<uint64_t>(b) * kPrepBlock + t;
  uint64_t key = 0;
  bool e = false;
  if (idx < n) {
    key = R[idx];
    e = edge_exists(g, key_src(key), key_dst(key));
  }
  const uint64_t m = __ballot(e);
  const uint32_t r = static_cast<uint32_t>(__popcll(m & lanemask_lt()));
  if ((t & 63) == 0) wcnt[w] = static_cast<uint32_t>(__popcll(m));
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (uint32_t v = 0; v < kPrepBlock / 64; v++) {
    before += v < w ? wcnt[v] : 0u;
    tot += wcnt[v];
  }
  if (t < 64) {  // wave 0 walks the look-back
    const uint32_t off = lbk::lookback_wave(state, 1, b, 0, err, epoch, tot);
    if (t == 0) {
      s_off = off;
      if (b == nblocks - 1) *nout = off + tot;
    }
  }
  __syncthreads();
  if (e) out[s_off + before + r] = key;
}

__global__ void __launch_bounds__(kPrepBlock) k_delete_prep(GView g, const uint64_t *R, const uint64_t *np,
                                                            uint64_t *out, uint64_t *nout, uint64_t *state,
                                                            uint32_t *misc, uint32_t *err, uint32_t nblocks,
                                                            uint32_t epoch) {
  __shared__ uint32_t s_ticket;
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  delete_prep_block(b, g, R, np, out, nout, state, err, nblocks, epoch);
}

// Both halves of a mixed batch in one launch, so the two latency-bound
// lookup passes overlap: blocks in start order [0, nb) prepare the inserted
// keys I (look-back words state[0, 2 nb)), blocks [nb, 2 nb) the deleted
// keys D (words state[2 nb, 4 nb)).  A block of either half only waits on
// blocks of its half that started before it.  The deleted keys get the
// same outputs as the inserted ones: existence flags (dex), the prefix of
// the absent keys (dnidx; the existing keys before i are i - dnidx[i]) and
// their source runs (drun_beg, count *druns), so the delete half runs over
// D itself and no compaction is needed.  ileaf: the PMA leaf each inserted
// key's lookup searched (~0 when it did not reach the PMA); dslot: the PMA
// slot of each deleted key (~0 when it is not in the PMA), a hint for the
// PMA delete.
__global__ void __launch_bounds__(kPrepBlock) k_mixed_prep(GView g, const uint64_t *I, const uint64_t *ni,
                                                           uint64_t cap, uint32_t *ex, uint32_t *nidx,
                                                           uint32_t *run_beg, uint64_t *nruns, uint32_t *ileaf,
                                                           const uint64_t *D,
                                                           const uint64_t *nd, uint32_t *dex, uint32_t *dnidx,
                                                           uint32_t *drun_beg, uint64_t *druns, uint64_t *dslot,
                                                           uint64_t *state,
                                                           uint32_t *misc, uint32_t *err, uint32_t nb,
                                                           uint32_t epoch) {
  __shared__ uint32_t s_ticket;
  const uint32_t b = lbk::ticket(misc, err, 2 * nb, &s_ticket);
  if (b >= 2 * nb) return;
  if (b < nb)
    insert_prep_block<1, true>(b, g, I, ni, cap, ex, nidx, run_beg, nruns, state, err, nb, epoch, ileaf);
  else
    insert_prep_block<1, true>(b - nb, g, D, nd, cap, dex, dnidx, drun_beg, druns, state + 2ull * nb, err, nb,
                               epoch, nullptr, dslot);
}

// ------------------------------------------------------------ gates
// A batch's update kernels are launched before the host has read how the
// preparation went.  The first kernel of each half decides from the batch
// scalars whether the half runs and publishes the key and run counts the
// half's later kernels use: zero when it does not run, so such a half
// writes nothing to the graph.
//
// The preparation raised no error: no id out of range, no bucket beyond its
// capacity, no look-back wait given up.
__device__ inline bool prep_clean(const uint64_t *dsc) {
  return (dsc[0] & (1ull | kSortOverflow)) == 0 && dsc[bsw::kErrSort] == 0 && dsc[bsw::kErrScan] == 0;
}

// The insert half of the same batch completed on the device: no promotion
// pending, the in-place tree update done (or no tree keys), the in-leaf PMA
// merge done (or no PMA keys).
__device__ inline bool insert_half_done(const uint64_t *dsc, bool pma_launched) {
  return dsc[5] == 0 && (dsc[4] == 0 || dsc[13] == 0) && (dsc[3] == 0 || (pma_launched && dsc[7] == 0));
}

// The counts of a half: cw[0] keys, cw[1] runs (at most the capacities)
// when ok, else zero; block 0 publishes them with the outcome at gw.
struct HalfCounts {
  uint64_t keys;
  uint32_t runs;
};
__device__ inline HalfCounts gate_counts(bool ok, const uint64_t *cw, uint64_t cap_keys, uint32_t cap_runs,
                                         uint64_t *gw) {
  HalfCounts c{0, 0};
  if (ok) {
    c.keys = dev_count(cw, cap_keys);
    c.runs = static_cast<uint32_t>(dev_count(cw + 1, cap_runs));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    gw[0] = c.keys;
    gw[1] = c.runs;
    gw[2] = ok ? 1u : 2u;
  }
  return c;
}

// ---- Per-run kernels with their own offsets
// The first kernel of each half gives one thread to each run (256 runs per
// block) and computes, besides the run's update, the run's share of the
// half's output lists (keys for the PMA, for the B-tree, ...).  Their
// exclusive prefixes over the launch come from a block scan and the
// launch's decoupled look-back (one column per list), so the kernel writes
// its lists in run order itself: no flag array, no separate select or scan.
constexpr uint32_t kRunBlock = 256;

// C <= 4 columns (one look-back wave each).  pre[c]: this thread's
// exclusive prefix over the launch; tot[c]: the prefix through this block
// (the grand total in the last block).  Called by every thread of the block.
template <int C>
__device__ inline void run_block_scan(const uint32_t (&x)[C], uint32_t (&pre)[C], uint32_t (&tot)[C], uint64_t *state,
                                      uint32_t b, uint32_t *err, uint32_t epoch) {
  static_assert(C >= 1 && C <= static_cast<int>(kRunBlock / 64), "one look-back wave per column");
  __shared__ uint32_t s_w[C][kRunBlock / 64];
  __shared__ uint32_t s_off[C], s_agg[C];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc[C];
#pragma unroll
  for (int c = 0; c < C; c++) {
    uint32_t v = x[c];
    for (uint32_t o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(v, o, 64);
      if (lane >= o) v += y;
    }
    inc[c] = v;
    if (lane == 63) s_w[c][w] = v;
  }
  __syncthreads();
  if (w < static_cast<uint32_t>(C)) {  // wave c walks column c
    uint32_t agg = 0;
#pragma unroll
    for (uint32_t i = 0; i < kRunBlock / 64; i++) agg += s_w[w][i];
    const uint32_t off = lbk::lookback_wave(state, C, b, w, err, epoch, agg);
    if (lane == 0) {
      s_off[w] = off;
      s_agg[w] = agg;
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < C; c++) {
    uint32_t base = s_off[c];
    for (uint32_t i = 0; i < w; i++) base += s_w[c][i];
    pre[c] = base + inc[c] - x[c];
    tot[c] = s_off[c] + s_agg[c];
  }
}

// ------------------------------------------------------------ inline merge
// One thread per source run of the unique sorted batch R.  ex[] marks keys
// already present; nidx[] is the exclusive count of new keys (size |R|+1).
// Writes the new inline slots and degree, the run's route, and the old
// inline entries that are kicked out (kick_old[run*13 ..], count in
// run_nko).
// k_inline_merge in registers for a run of at most kMergeRun keys, or a
// longer run whose first kMergeRun keys hold at least kInline new keys
// (then only they can enter the inline slots; any later key is larger).
// The window's keys and new-key flags are loaded at once, and every inline
// entry and new key computes its rank in the merged order (inline and new
// keys are disjoint); ranks below kInline are the new inline slots, the
// old entries ranked beyond are kicked out (the largest old ones).  Same
// outputs as the general loop; returns false, having written nothing, for
// a run it does not cover.
constexpr uint32_t kMergeRun = 16;

// What the merge decided for one run.
struct RunMerge {
  uint8_t route = kRouteNone;
  uint32_t c = 0;     // new keys that entered the inline slots
  uint32_t nko = 0;   // old inline entries pushed out (into ko[0, nko), sorted)
  uint32_t nnew = 0;  // new keys of the run
  uint32_t promote = 0;
};

__device__ inline bool inline_merge_window(GView g, const uint64_t *R, const uint32_t *ex, const uint32_t *nidx,
                                           uint32_t b0, uint32_t b1, uint32_t s, RunMerge &m, uint32_t *ko,
                                           uint32_t &old_degree) {
  const uint32_t L = b1 - b0;
  uint32_t rd[kMergeRun];
  bool rnew[kMergeRun];
#pragma unroll
  for (uint32_t j = 0; j < kMergeRun; j++) {
    rd[j] = j < L ? key_dst(R[b0 + j]) : 0u;
    rnew[j] = j < L && ex[b0 + j] == 0;
  }
  uint32_t wnew = 0;
#pragma unroll
  for (uint32_t j = 0; j < kMergeRun; j++) wnew += rnew[j] ? 1u : 0u;
  if (L > kMergeRun && wnew < kInline) return false;
  const uint4 *vp = reinterpret_cast<const uint4 *>(g.vb + s);
  const uint4 q0 = vp[0], q1 = vp[1], q2 = vp[2], q3 = vp[3];
  const uint32_t w[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                          q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
  const uint32_t D = w[0];
  old_degree = D;  // promotions need the degree before the merge
  const uint32_t lim = D < kInline ? D : kInline;
  const bool is_tree = (w[15] & 0x80000000u) != 0;  // aux bit 63
  const uint32_t nnew = L > kMergeRun ? nidx[b1] - nidx[b0] : wnew;
  uint32_t dup_inline = 0;
#pragma unroll
  for (uint32_t j = 0; j < kMergeRun; j++) {
    bool in = false;
#pragma unroll
    for (uint32_t p = 0; p < kInline; p++) in |= p < lim && w[1 + p] == rd[j];
    dup_inline += (j < L && !rnew[j] && in) ? 1u : 0u;
  }
  // k = |inline ∪ run| as the reference counts it (graph.h:290-330); a run
  // longer than kMergeRun overflows whatever the duplicates (k >= |run|)
  const uint64_t size = L;
  const uint64_t kk = static_cast<uint64_t>(lim) + size - dup_inline;
  const bool overflow = L > kMergeRun || ((D == 0) ? (size > kInline) : (kk > kInline));
  uint8_t route = kRouteNone;
  if (overflow) route = (is_tree || static_cast<uint64_t>(D) + size > kMediumDegree) ? kRouteTree : kRoutePma;
  VBlock &vb = g.vb[s];
  // old inline entries: rank = p + new keys below; the kept ones are a prefix
  uint32_t rank_old[kInline];
  uint32_t kept_old = 0;
#pragma unroll
  for (uint32_t p = 0; p < kInline; p++) {
    uint32_t rk = p;
#pragma unroll
    for (uint32_t j = 0; j < kMergeRun; j++) rk += (rnew[j] && rd[j] < w[1 + p]) ? 1u : 0u;
    rank_old[p] = rk;
    kept_old += (p < lim && rk < kInline) ? 1u : 0u;
  }
#pragma unroll
  for (uint32_t p = 0; p < kInline; p++) {
    if (p >= lim) continue;
    if (rank_old[p] < kInline) {
      if (rank_old[p] != p) vb.nbr[rank_old[p]] = w[1 + p];
    } else {
      ko[p - kept_old] = w[1 + p];
    }
  }
  // new keys: rank = inline entries below + new keys before it in the run
  uint32_t c = 0, before = 0;
#pragma unroll
  for (uint32_t j = 0; j < kMergeRun; j++) {
    if (!rnew[j]) continue;
    uint32_t rk = before;
#pragma unroll
    for (uint32_t p = 0; p < kInline; p++) rk += (p < lim && w[1 + p] < rd[j]) ? 1u : 0u;
    if (rk < kInline) {
      vb.nbr[rk] = rd[j];
      c++;
    }
    before++;
  }
  vb.degree = D + nnew;
  m.route = route;
  m.c = c;
  m.nko = lim - kept_old;
  m.nnew = nnew;
  m.promote = (route == kRouteTree && !is_tree) ? 1u : 0u;
  return true;
}

// The serial merge for the runs the window does not cover.
__device__ inline void inline_merge_serial(GView g, const uint64_t *R, const uint32_t *ex, const uint32_t *nidx,
                                           uint32_t b0, uint32_t b1, uint32_t s, RunMerge &m, uint32_t *ko,
                                           uint32_t &old_degree) {
  VBlock &vb = g.vb[s];
  uint32_t D = vb.degree;
  old_degree = D;  // promotions need the degree before the merge
  uint32_t lim = D < kInline ? D : kInline;
  uint32_t inl[kInline];
  for (uint32_t i = 0; i < kInline; i++) inl[i] = i < lim ? vb.nbr[i] : 0;
  uint64_t size = b1 - b0;  // run length including keys already present
  uint32_t nnew = nidx[b1] - nidx[b0];
  // k = |inline ∪ run| as the reference counts it (graph.h:290-330)
  uint32_t dup_inline = 0;
  for (uint32_t p = 0; p < lim; p++) {
    uint64_t x = make_key(s, inl[p]);
    uint64_t lb = b0 + lower_bound_dev(R + b0, b1 - b0, x);
    if (lb < b1 && R[lb] == x) dup_inline++;
  }
  uint64_t kk = static_cast<uint64_t>(lim) + size - dup_inline;
  bool is_tree = (vb.aux & kAuxTree) != 0;
  bool overflow = (D == 0) ? (size > kInline) : (kk > kInline);
  uint8_t route = kRouteNone;
  if (overflow) route = (is_tree || static_cast<uint64_t>(D) + size > kMediumDegree) ? kRouteTree : kRoutePma;
  // merge the inline slots with the new keys, keep the 13 smallest
  uint32_t out[kInline];
  uint32_t ni = 0, p = 0, c = 0;
  uint32_t q = b0;
  while (ni < kInline) {
    while (q < b1 && ex[q]) q++;
    bool have_a = p < lim, have_b = q < b1;
    if (!have_a && !have_b) break;
    if (have_a && (!have_b || inl[p] < key_dst(R[q]))) {
      out[ni++] = inl[p++];
    } else {
      out[ni++] = key_dst(R[q]);
      q++;
      c++;
    }
  }
  uint32_t nko = lim - p;
  for (uint32_t i = 0; i < nko; i++) ko[i] = inl[p + i];
  for (uint32_t i = 0; i < ni; i++) vb.nbr[i] = out[i];
  vb.degree = D + nnew;
  m.route = route;
  m.c = c;
  m.nko = nko;
  m.nnew = nnew;
  m.promote = (route == kRouteTree && !is_tree) ? 1u : 0u;
}

// The insert half's first kernel, one thread per run of the unique sorted
// batch R (ex[] marks keys already present, nidx[] the exclusive count of
// new keys): its gate (gate null: open) and counts (cw, published at gw);
// the run's inline merge (new inline slots and degree, its route); then the
// keys the run kicks out of the inline tier: the new keys beyond the first
// c, merged with the old inline entries pushed out, written sorted to the
// run's offset in the PMA list P or the B-tree list T.  The offsets and the
// promotion prefix (ppos) come from the run scan; the block's keys are
// placed by all its threads (a run's keys are contiguous in R).  Each PMA
// key also gets the leaf its preparation lookup searched (Pleaf, from
// ileaf when given; ~0 when unknown, for k_leaf_of to search; the PMA does
// not change before the in-leaf merge reads it).  Totals: sc[0] PMA keys, sc[1] tree keys,
// sc[2] promotions.
__global__ void __launch_bounds__(kRunBlock) k_insert_runs(GView g, const uint64_t *R, const uint32_t *ex,
                                                           const uint32_t *nidx, const uint32_t *ileaf,
                                                           const uint32_t *run_beg,
                                                           const uint64_t *cw, const uint64_t *gate, uint64_t *gw,
                                                           uint64_t cap_keys, uint32_t cap_runs, uint64_t *P,
                                                           uint32_t *Pleaf, uint64_t *T, uint32_t *promote,
                                                           uint32_t *ppos,
                                                           uint32_t *old_deg, uint64_t *sc, uint64_t *state,
                                                           uint32_t *misc, uint32_t *err, uint32_t nblocks,
                                                           uint32_t epoch) {
  __shared__ uint32_t s_ticket;
  __shared__ uint32_t s_b0[kRunBlock + 1], s_nb0[kRunBlock], s_c[kRunBlock], s_off[kRunBlock];
  __shared__ uint8_t s_route[kRunBlock], s_nko[kRunBlock];
  __shared__ uint32_t s_ko[kRunBlock][kInline];
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  const HalfCounts hc = gate_counts(!gate || prep_clean(gate), cw, cap_keys, cap_runs, gw);
  const uint32_t t = threadIdx.x;

  const uint64_t r0 = static_cast<uint64_t>(b) * kRunBlock;
  const uint32_t nlive = r0 < hc.runs ? static_cast<uint32_t>(hc.runs - r0 < kRunBlock ? hc.runs - r0 : kRunBlock) : 0u;
  const bool live = t < nlive;
  const uint64_t r = r0 + t;
  RunMerge m;
  uint32_t b0 = 0, b1 = 0, nb0 = 0;
  if (live) {
    b0 = run_beg[r];
    b1 = run_beg[r + 1];
    nb0 = nidx[b0];
    const uint32_t s = key_src(R[b0]);
    uint32_t od = 0;
    if (!inline_merge_window(g, R, ex, nidx, b0, b1, s, m, s_ko[t], od))
      inline_merge_serial(g, R, ex, nidx, b0, b1, s, m, s_ko[t], od);
    old_deg[r] = od;
  }
  const uint32_t K = m.route != kRouteNone ? m.nnew - m.c + m.nko : 0u;  // keys kicked out of the inline tier
  const uint32_t x[3] = {m.route == kRoutePma ? K : 0u, m.route == kRouteTree ? K : 0u, m.promote};
  uint32_t pre[3], tot[3];
  run_block_scan<3>(x, pre, tot, state, b, err, epoch);
  if (live) {
    promote[r] = m.promote;
    ppos[r] = pre[2];
  }
  if (live) s_b0[t] = b0;  // s_b0[nlive] is the end (thread 0, below)
  s_nb0[t] = nb0;
  s_c[t] = m.c;
  s_off[t] = m.route == kRoutePma ? pre[0] : pre[1];
  s_route[t] = m.route;
  s_nko[t] = static_cast<uint8_t>(m.nko);
  if (t == 0) s_b0[nlive] = nlive ? run_beg[r0 + nlive] : 0u;  // run_beg[nruns] = nR
  __syncthreads();
  if (b == nblocks - 1 && t == 0) {
    sc[0] = tot[0];
    sc[1] = tot[1];
    sc[2] = tot[2];
  }
  if (nlive == 0) return;
  // the old inline entries each run pushed out, at their ranks among its
  // kicked keys (new tail keys below them + their own order)
  if (live && m.route != kRouteNone) {
    const bool pma = m.route == kRoutePma;
    const uint64_t o = s_off[t];
    const uint32_t s = key_src(R[b0]);
    for (uint32_t e = 0; e < m.nko; e++) {
      const uint32_t xo = s_ko[t][e];
      const uint64_t key = make_key(s, xo);
      const uint64_t lb = b0 + lower_bound_dev(R + b0, b1 - b0, key);
      const uint32_t new_less = nidx[lb] - nb0 - m.c;
      if (pma) {
        P[o + e + new_less] = key;
        Pleaf[o + e + new_less] = ~0u;  // searched by k_leaf_of
      } else {
        T[o + e + new_less] = key;
      }
    }
  }
  // the new keys beyond the first c of each run, by every thread over the
  // block's keys: rank among the new keys, plus the pushed-out entries below
  const uint32_t kb = s_b0[0], ke = s_b0[nlive];
  for (uint32_t i = kb + t; i < ke; i += kRunBlock) {
    if (ex[i]) continue;
    uint32_t lo = 0, hi = nlive;  // run: last q with s_b0[q] <= i
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_b0[mid] <= i) lo = mid;
      else hi = mid;
    }
    const uint8_t route = s_route[lo];
    if (route == kRouteNone) continue;
    const uint32_t j = nidx[i] - s_nb0[lo];
    const uint32_t c = s_c[lo];
    if (j < c) continue;
    const uint64_t key = R[i];
    const uint32_t d = key_dst(key);
    const uint32_t nko = s_nko[lo];
    uint32_t less = 0;
    while (less < nko && s_ko[lo][less] < d) less++;
    const uint64_t pos = s_off[lo] + (j - c) + less;
    if (route == kRoutePma) {
      P[pos] = key;
      Pleaf[pos] = ileaf ? ileaf[i] : ~0u;  // ~0: searched by k_leaf_of
    } else {
      T[pos] = key;
    }
  }
}

// ------------------------------------------------------------ PMA kernels
// The leaf of each key; with only_unknown, only of the keys whose leaf
// reads ~0 (the others came from the preparation's lookups).
__global__ void k_leaf_of(GView g, const uint64_t *keys, uint64_t cap, uint32_t *leaf, uint64_t *fit_flag,
                          const uint64_t *kdev, bool only_unknown) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *fit_flag = 0;  // set later by k_pma_leaf_fit
  const uint64_t k = dev_count(kdev, cap);
  GRID_STRIDE(i, k) {
    if (only_unknown && leaf[i] != ~0u) continue;
    const uint64_t key = keys[i];
    leaf[i] = static_cast<uint32_t>(pma_leaf_of(g, key, g.vb[key_src(key)].aux));
  }
}

__global__ void k_climb(GView g, Bounds bd, const uint32_t *T, const uint32_t *Tstart, const uint64_t *ntp,
                        uint8_t *mark, uint32_t *chosen) {
  const uint32_t nt = static_cast<uint32_t>(*ntp);
  GRID_STRIDE(t, nt) {
    uint32_t f = T[t];
    uint32_t L = g.H;
    for (uint32_t l = 0; l <= g.H; l++) {
      uint64_t lo = (static_cast<uint64_t>(f) >> l) << l;
      uint64_t hi = lo + (1ull << l);
      uint64_t a = lower_bound_dev<uint32_t>(T, nt, static_cast<uint32_t>(lo));
      uint64_t b = hi >= (1ull << g.H) ? nt : lower_bound_dev<uint32_t>(T, nt, static_cast<uint32_t>(hi));
      uint64_t newc = Tstart[b] - Tstart[a];
      uint64_t old = g.pcnt[tree_level_off(g.H, l) + (f >> l)];
      if (old + newc <= bd.up[l]) {
        L = l;
        break;
      }
    }
    mark[tree_level_off(g.H, L) + (f >> L)] = 1;
    chosen[t] = L;
  }
}

__global__ void k_cover(GView g, const uint32_t *T, const uint64_t *ntp, const uint8_t *mark, const uint32_t *chosen,
                        uint32_t *cover) {
  const uint32_t nt = static_cast<uint32_t>(*ntp);
  GRID_STRIDE(t, nt) {
    uint32_t f = T[t];
    uint32_t L = chosen[t];
    for (uint32_t l = g.H; l > L; l--) {
      if (mark[tree_level_off(g.H, l) + (f >> l)]) {
        L = l;
        break;
      }
    }
    cover[t] = L;
  }
}

// Deletes: the window of each underfull leaf T[t] is its lowest ancestor
// whose count is above that level's lower bound (the climb of
// src/PMA.cc:1646-1712); the root when none is (an array already at its
// smallest size).
__global__ void k_climb_low(GView g, Bounds bd, const uint32_t *T, const uint64_t *ntp, uint8_t *mark,
                            uint32_t *chosen) {
  const uint32_t nt = static_cast<uint32_t>(*ntp);
  GRID_STRIDE(t, nt) {
    const uint32_t f = T[t];
    uint32_t L = g.H;
    for (uint32_t l = 1; l < g.H; l++) {
      if (g.pcnt[tree_level_off(g.H, l) + (f >> l)] > bd.lo[l]) {
        L = l;
        break;
      }
    }
    mark[tree_level_off(g.H, L) + (f >> L)] = 1;
    chosen[t] = L;
  }
}

// Window record: lo leaf, level, key range, old element count.
struct Win {
  uint64_t lo;
  uint32_t L;
  uint32_t kb, ke;
  uint32_t old;
};

__global__ void k_clear_marks(GView g, const uint32_t *T, const uint64_t *ntp, const uint32_t *chosen, uint8_t *mark) {
  const uint32_t nt = static_cast<uint32_t>(*ntp);
  GRID_STRIDE(t, nt) {
    uint32_t L = chosen[t];
    mark[tree_level_off(g.H, L) + (T[t] >> L)] = 0;
  }
}

// ---- PMA inserts that fit their leaves
// When every touched leaf can take its new keys and stay within the leaf
// density bound (old + new <= up[0]), the window search picks each leaf as
// its own window and the redistribution is a sorted merge inside the leaf.
// Two kernels produce that same state directly: k_pma_leaf_fit checks the
// bound, and k_pma_leaf_merge writes the merged leaf, its count and fence
// and the run-start begin pointers (as k_gather and k_fill do between
// them).  The merge does nothing when a leaf does not fit; the window path
// then runs.

// Also closes Tstart (Tstart[nt] = k) for the kernels after it.
__global__ void k_pma_leaf_fit(GView g, uint64_t up0, const uint32_t *T, uint32_t *Tstart, const uint64_t *ntp,
                               uint32_t cap, uint64_t *flag, const uint64_t *kdev) {
  const uint32_t k = static_cast<uint32_t>(dev_count(kdev, cap));
  const uint32_t nt = static_cast<uint32_t>(*ntp);
  if (blockIdx.x == 0 && threadIdx.x == 0) Tstart[nt] = k;
  GRID_STRIDE(t, nt) {
    const uint64_t newc = (t + 1 < nt ? Tstart[t + 1] : k) - Tstart[t];
    if (g.pcnt[T[t]] + newc > up0) atomicOr(reinterpret_cast<unsigned long long *>(flag), 1ull);
  }
}


// kLeafLanes lanes per touched leaf (the grid covers every possible leaf;
// groups past the touched count exit).  Lane l of a group reads slots
// kLaneSlots * l .. + kLaneSlots - 1 (16-byte loads); the leaf's elements
// (a leaf may hold holes after deletes) are compacted by a group prefix
// into LDS next to the leaf's new keys; then lane l finds the merge-path
// split for its output slots, merges its outputs and writes them back with
// 16-byte stores, packed left, the rest of the leaf emptied.  16 lanes (four
// leaves per wave) measured fastest: one wave per leaf needs four times the
// waves, and 8 lanes per leaf doubles the LDS per workgroup (lower
// occupancy) and the serial merge per lane.
constexpr uint32_t kLeafLanes = 16, kLaneSlots = kLeafSlots / kLeafLanes, kLaneQuads = kLaneSlots / 4;
constexpr uint32_t kMergeGroups = 256 / kLeafLanes;  // leaves per 256-thread workgroup

__global__ void __launch_bounds__(256) k_pma_leaf_merge(GView g, const uint32_t *T, const uint32_t *Tstart,
                                                        const uint64_t *ntp, const uint64_t *keys,
                                                        const uint64_t *flag, uint64_t *err) {
  if (*flag) return;
  const uint32_t nt = static_cast<uint32_t>(*ntp);
  if (static_cast<uint64_t>(blockIdx.x) * kMergeGroups >= nt) return;  // the whole workgroup is idle
  __shared__ uint64_t s_old[kMergeGroups][kLeafSlots];
  __shared__ uint64_t s_new[kMergeGroups][kLeafSlots];
  constexpr uint64_t kNone = ~0ull;
  const uint32_t q = threadIdx.x / kLeafLanes, l = threadIdx.x % kLeafLanes;
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kMergeGroups + q;
  const bool live = t < nt;
  uint64_t leaf = 0, base = 0;
  uint32_t nb = 0, c = 0;
  uint32_t sv[kLaneSlots] = {};  // old sources of this lane's slots (kept in emptied slots)
  uint32_t dd[kLaneSlots] = {};  // old destinations (a quad that stays the same is not stored)
  if (live) {
    leaf = T[t];
    base = leaf << kLeafShift;
    const uint32_t kb = Tstart[t];
    nb = Tstart[t + 1] - kb;
#pragma unroll
    for (uint32_t h = 0; h < kLaneQuads; h++) {
      const uint4 dv = reinterpret_cast<const uint4 *>(g.pdst + base)[kLaneQuads * l + h];
      const uint4 sq = reinterpret_cast<const uint4 *>(g.psrc + base)[kLaneQuads * l + h];
      dd[4 * h] = dv.x, dd[4 * h + 1] = dv.y, dd[4 * h + 2] = dv.z, dd[4 * h + 3] = dv.w;
      sv[4 * h] = sq.x, sv[4 * h + 1] = sq.y, sv[4 * h + 2] = sq.z, sv[4 * h + 3] = sq.w;
    }
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t j = 0; j < kLaneSlots; j++) cnt += dd[j] != kEmpty ? 1u : 0u;
    uint32_t inc = cnt;
    for (uint32_t o = 1; o < kLeafLanes; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, kLeafLanes);
      if (l >= o) inc += y;
    }
    c = __shfl(inc, kLeafLanes - 1, kLeafLanes);
    uint32_t pos = inc - cnt;
#pragma unroll
    for (uint32_t j = 0; j < kLaneSlots; j++)
      if (dd[j] != kEmpty) s_old[q][pos++] = make_key(sv[j], dd[j]);
    if (l == 0 && c != g.pcnt[leaf]) atomicOr(reinterpret_cast<unsigned long long *>(err), 2ull);
    for (uint32_t j = l; j < nb; j += kLeafLanes) s_new[q][j] = keys[kb + j];
  }
  __syncthreads();
  if (!live) return;
  // c + nb <= up[0] < 64.  Merge-path split for output o0: i old elements
  // and o0 - i new keys come before it (the two lists are disjoint).
  const uint64_t *A = s_old[q], *B = s_new[q];
  const uint32_t total = c + nb, o0 = kLaneSlots * l;
  uint32_t lo = o0 > nb ? o0 - nb : 0, hi = o0 < c ? o0 : c;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (A[mid] < B[o0 - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  uint32_t i = lo, j = o0 - lo;
  // the output before o0: the larger of the last old and last new taken
  uint64_t prev = kNone;
  if (o0 > 0 && o0 <= total) {
    const uint64_t pa = i > 0 ? A[i - 1] : 0, pb = j > 0 ? B[j - 1] : 0;
    prev = pa > pb ? pa : pb;
  }
  uint64_t out[kLaneSlots];
#pragma unroll
  for (uint32_t r = 0; r < kLaneSlots; r++) {
    out[r] = kNone;
    if (o0 + r < total) {
      if (j >= nb || (i < c && A[i] < B[j])) out[r] = A[i++];
      else out[r] = B[j++];
    }
  }
#pragma unroll
  for (uint32_t h = 0; h < kLaneQuads; h++) {
    uint4 nd, ns;
    const uint32_t r = 4 * h;
    nd.x = o0 + r < total ? key_dst(out[r]) : kEmpty;
    nd.y = o0 + r + 1 < total ? key_dst(out[r + 1]) : kEmpty;
    nd.z = o0 + r + 2 < total ? key_dst(out[r + 2]) : kEmpty;
    nd.w = o0 + r + 3 < total ? key_dst(out[r + 3]) : kEmpty;
    ns.x = o0 + r < total ? key_src(out[r]) : sv[r];
    ns.y = o0 + r + 1 < total ? key_src(out[r + 1]) : sv[r + 1];
    ns.z = o0 + r + 2 < total ? key_src(out[r + 2]) : sv[r + 2];
    ns.w = o0 + r + 3 < total ? key_src(out[r + 3]) : sv[r + 3];
    // the slots before the first new key keep their elements (unless holes
    // left by deletes shift them): only quads that change are written
    if (nd.x != dd[r] || nd.y != dd[r + 1] || nd.z != dd[r + 2] || nd.w != dd[r + 3])
      reinterpret_cast<uint4 *>(g.pdst + base)[kLaneQuads * l + h] = nd;
    if (ns.x != sv[r] || ns.y != sv[r + 1] || ns.z != sv[r + 2] || ns.w != sv[r + 3])
      reinterpret_cast<uint4 *>(g.psrc + base)[kLaneQuads * l + h] = ns;
  }
  // Run starts in this leaf move the vertex's begin pointer.  A pointer
  // into this leaf (the run started here) moves with a CAS, which fails
  // only when an earlier leaf's wave has already lowered it to a new
  // earlier element; any other pointer takes the minimum.  Every
  // interleaving leaves the slot of the vertex's first element.  A lane's
  // pointer loads are issued together: each of its outputs that starts a
  // run belongs to a different vertex.
  bool start[kLaneSlots];
  unsigned long long cur[kLaneSlots];
#pragma unroll
  for (uint32_t r = 0; r < kLaneSlots; r++) {
    const uint64_t before = r == 0 ? prev : out[r - 1];
    start[r] = o0 + r < total && (before == kNone || key_src(before) != key_src(out[r]));
    cur[r] = start[r] ? __hip_atomic_load(reinterpret_cast<unsigned long long *>(&g.vb[key_src(out[r])].aux),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      : 0ull;
  }
#pragma unroll
  for (uint32_t r = 0; r < kLaneSlots; r++) {
    if (!start[r]) continue;
    unsigned long long *aux = reinterpret_cast<unsigned long long *>(&g.vb[key_src(out[r])].aux);
    const unsigned long long slot = base + o0 + r;
    if (cur[r] >= base && cur[r] < base + kLeafSlots) atomicCAS(aux, cur[r], slot);
    else atomicMin(aux, slot);
  }
  if (l == 0) {
    g.pcnt[leaf] = total;
    g.fence[leaf] = out[0];  // total >= 1: the leaf's first key
  }
}

// Tiny helpers that move device counts around without a host round trip.
__global__ void k_put_u32(uint32_t *p, const uint64_t *idx, uint32_t v) { p[*idx] = v; }
__global__ void k_u32_to_u64(const uint32_t *a, uint64_t *b) { *b = *a; }
__global__ void k_u32_to_u64_n(const uint32_t *a, uint64_t n, uint64_t *b) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    b[i] = a[i];
}
__global__ void k_u32_pair(const uint32_t *a, const uint32_t *b, uint64_t *out) {
  out[0] = *a;
  out[1] = *b;
}

// ---- selectors (select.hpp): predicate + emit, no flag arrays
// first key of each run of equal keys, with its weight
struct SelUniqueW {
  const uint64_t *k;
  const uint32_t *w;
  uint64_t *out;
  uint32_t *wout;
  __device__ bool pred(uint64_t i) const { return i == 0 || k[i] != k[i - 1]; }
  __device__ void emit(uint64_t i, uint64_t pos) const {
    out[pos] = k[i];
    wout[pos] = w[i];
  }
};
struct SelUnique {
  const uint64_t *k;
  uint64_t *out;
  __device__ bool pred(uint64_t i) const { return i == 0 || k[i] != k[i - 1]; }
  __device__ void emit(uint64_t i, uint64_t pos) const { out[pos] = k[i]; }
};
// Sorted tagged keys: the last key of each edge is its final op; inserted
// edges go to output 1 (I), deleted ones to output 2 (D), untagged.
struct SelSplitOps {
  const uint64_t *t;
  uint64_t n;
  uint64_t *I, *D;
  __device__ uint32_t pick(uint64_t i) const {
    const bool last = i + 1 == n || (t[i] >> 1) != (t[i + 1] >> 1);
    return last ? 1u + static_cast<uint32_t>(t[i] & 1) : 0u;
  }
  __device__ void emit(uint64_t i, uint32_t which, uint64_t pos) const { (which == 1 ? I : D)[pos] = untag(t[i]); }
};
struct SelRunStart {
  const uint64_t *k;
  uint32_t *beg;
  __device__ bool pred(uint64_t i) const { return i == 0 || key_src(k[i]) != key_src(k[i - 1]); }
  __device__ void emit(uint64_t i, uint64_t pos) const { beg[pos] = static_cast<uint32_t>(i); }
};
// run starts among the first *np keys
struct SelRunStartN {
  const uint64_t *k;
  uint32_t *beg;
  const uint64_t *np;
  __device__ bool pred(uint64_t i) const { return i < *np && (i == 0 || key_src(k[i]) != key_src(k[i - 1])); }
  __device__ void emit(uint64_t i, uint64_t pos) const { beg[pos] = static_cast<uint32_t>(i); }
  __device__ void finish(uint64_t runs) const { beg[runs] = static_cast<uint32_t>(*np); }  // beg[nruns] = n
};
struct SelFlagged {
  const uint64_t *k;
  const uint32_t *flag;
  uint64_t *out;
  __device__ bool pred(uint64_t i) const { return flag[i] != 0; }
  __device__ void emit(uint64_t i, uint64_t pos) const { out[pos] = k[i]; }
};
struct SelExisting {
  GView g;
  const uint64_t *k;
  uint64_t *out;
  __device__ bool pred(uint64_t i) const { return edge_exists(g, key_src(k[i]), key_dst(k[i])); }
  __device__ void emit(uint64_t i, uint64_t pos) const { out[pos] = k[i]; }
};
// Both tier lists in one pass over the vertex blocks: 1 = PMA (more than
// the inline slots and no tree), 2 = B-tree.
struct SelTiers {
  const VBlock *vb;
  uint32_t *pma, *tree;
  __device__ uint32_t pick(uint64_t i) const {
    const uint64_t aux = vb[i].aux;
    if (aux & kAuxTree) return 2u;
    return vb[i].degree > kInline ? 1u : 0u;
  }
  __device__ void emit(uint64_t i, uint32_t which, uint64_t pos) const {
    (which == 1 ? pma : tree)[pos] = static_cast<uint32_t>(i);
  }
};
// Keys of the two secondary tiers: 1 = PMA (into P), 2 = B-tree (into T).
struct SelTouched {
  const uint32_t *leaf;
  uint32_t *T, *Tstart;
  const uint64_t *np = nullptr;  // keys past *np are not listed (null: all)
  __device__ bool pred(uint64_t i) const { return (!np || i < *np) && (i == 0 || leaf[i] != leaf[i - 1]); }
  __device__ void emit(uint64_t i, uint64_t pos) const {
    T[pos] = leaf[i];
    Tstart[pos] = static_cast<uint32_t>(i);
  }
};
// leaves holding at most lo0 elements (start offsets all 0: no new keys)
struct SelLowLeaf {
  const uint32_t *pcnt;
  uint32_t lo0;
  uint32_t *T, *Tstart;
  __device__ bool pred(uint64_t i) const { return pcnt[i] <= lo0; }
  __device__ void emit(uint64_t i, uint64_t pos) const {
    T[pos] = static_cast<uint32_t>(i);
    Tstart[pos] = 0;
  }
};
// last op of each key (keys sorted, arrival order kept inside a key)
struct SelLastOp {
  const uint64_t *k;
  const uint32_t *ops;
  uint64_t n;
  uint32_t want;
  uint64_t *out;
  __device__ bool pred(uint64_t i) const {
    bool last = (i + 1 == n) || k[i] != k[i + 1];
    return last && ((ops[i] != 0) == (want != 0));
  }
  __device__ void emit(uint64_t i, uint64_t pos) const { out[pos] = k[i]; }
};
// Window owner: the first touched leaf of each covering window; emits the
// window record and its leaf count.
struct SelOwner {
  GView g;
  const uint32_t *T, *Tstart, *cover;
  const uint64_t *ntp;
  Win *win;
  uint32_t *win_leaves;
  __device__ bool pred(uint64_t t) const {
    if (t >= *ntp) return false;
    uint32_t L = cover[t];
    return t == 0 || (static_cast<uint64_t>(T[t - 1]) >> L) != (static_cast<uint64_t>(T[t]) >> L);
  }
  __device__ void emit(uint64_t t, uint64_t pos) const {
    const uint32_t nt = static_cast<uint32_t>(*ntp);
    uint32_t L = cover[t];
    uint64_t lo = (static_cast<uint64_t>(T[t]) >> L) << L;
    uint64_t hi = lo + (1ull << L);
    uint64_t b = hi >= (1ull << g.H) ? nt : lower_bound_dev<uint32_t>(T, nt, static_cast<uint32_t>(hi));
    Win w;
    w.lo = lo;
    w.L = L;
    w.kb = Tstart[t];
    w.ke = Tstart[b];
    w.old = g.pcnt[tree_level_off(g.H, L) + (lo >> L)];
    win[pos] = w;
    win_leaves[pos] = 1u << L;
  }
};

// Window of leaf task j.  tw (task -> window map) is filled by
// k_task_map; without it, a binary search over task_base.
__device__ inline uint32_t task_window(const uint32_t *task_base, uint32_t nw, uint64_t j,
                                       const uint32_t *tw = nullptr) {
  if (tw) return tw[j];
  uint32_t lo = 0, hi = nw;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (task_base[mid] <= j) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

__global__ void k_task_map(const uint32_t *task_base, uint32_t nw, uint32_t *tw) {
  GRID_STRIDE(w, nw) {
    uint32_t b = task_base[w], e = task_base[w + 1];
    for (uint32_t j = b; j < e; j++) tw[j] = static_cast<uint32_t>(w);
  }
}

__global__ void k_task_counts(GView g, const Win *win, const uint32_t *task_base, uint32_t nw, uint64_t ntasks,
                              uint32_t *taskcnt, const uint32_t *tw) {
  GRID_STRIDE(j, ntasks) {
    uint32_t w = task_window(task_base, nw, j, tw);
    uint64_t leaf = win[w].lo + (j - task_base[w]);
    taskcnt[j] = g.pcnt[leaf];
  }
}

// One wave per leaf task: compact the leaf's elements into tmp_old and drop
// begin pointers that point into the window (they are re-derived by the
// fill kernel with atomicMin).
__global__ void __launch_bounds__(256) k_gather(GView g, const Win *win, const uint32_t *task_base, uint32_t nw,
                                                uint64_t ntasks, const uint32_t *taskoff, uint64_t *tmp_old,
                                                uint64_t *err, const uint32_t *tw) {
  uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint32_t lane = threadIdx.x & 63;
  for (uint64_t j = wave; j < ntasks; j += nwaves) {
    uint32_t w = task_window(task_base, nw, j, tw);
    uint64_t leaf = win[w].lo + (j - task_base[w]);
    uint64_t slot = (leaf << kLeafShift) + lane;
    uint32_t d = g.pdst[slot];
    bool v = d != kEmpty;
    uint64_t m = __ballot(v);
    if (lane == 0 && static_cast<uint32_t>(__popcll(m)) != g.pcnt[leaf])
      atomicOr(reinterpret_cast<unsigned long long *>(err), 2ull);
    if (v) {
      uint32_t s = g.psrc[slot];
      uint32_t r = __popcll(m & lanemask_lt());
      tmp_old[taskoff[j] + r] = make_key(s, d);
      if (g.vb[s].aux == slot) g.vb[s].aux = kPmaNone;
    }
  }
}

// One wave per leaf task: even redistribution of merge(old, new) over the
// window (count per leaf q or q+1, extras to the first leaves, packed
// left, as redistribute does in src/PMA.cc:419-544), fences, count tree.
__global__ void __launch_bounds__(256) k_fill(GView g, const Win *win, const uint32_t *task_base, uint32_t nw,
                                              uint64_t ntasks, const uint32_t *taskoff, const uint64_t *tmp_old,
                                              const uint64_t *keys, const uint32_t *tw) {
  uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint32_t lane = threadIdx.x & 63;
  for (uint64_t j = wave; j < ntasks; j += nwaves) {
    uint32_t w = task_window(task_base, nw, j, tw);
    Win W = win[w];
    uint64_t li = j - task_base[w];
    uint64_t leaf = W.lo + li;
    uint64_t na = W.old, nb = W.ke - W.kb;
    const uint64_t *A = tmp_old + taskoff[task_base[w]];
    const uint64_t *B = keys + W.kb;
    uint64_t tot = na + nb;
    uint64_t q = tot >> W.L, rem = tot & ((1ull << W.L) - 1);
    uint64_t c = q + (li < rem ? 1 : 0);
    uint64_t m0 = li * q + (li < rem ? li : rem);
    uint64_t slot = (leaf << kLeafShift) + lane;
    KeyArr AA{A}, BB{B};
    if (lane < c) {
      uint64_t m = m0 + lane;
      uint64_t i = merge_split(AA, na, BB, nb, m);
      uint64_t jb = m - i;
      uint64_t e = (i < na && (jb >= nb || A[i] < B[jb])) ? A[i] : B[jb];
      // the element before e is the larger of the last taken from A and B
      bool has_prev = m > 0;
      uint64_t prev = 0;
      if (i > 0) prev = A[i - 1];
      if (jb > 0 && (i == 0 || B[jb - 1] > prev)) prev = B[jb - 1];
      uint32_t s = key_src(e);
      g.pdst[slot] = key_dst(e);
      g.psrc[slot] = s;
      if (!has_prev || key_src(prev) != s)
        atomicMin(reinterpret_cast<unsigned long long *>(&g.vb[s].aux), static_cast<unsigned long long>(slot));
      if (lane == 0) g.fence[leaf] = e;
    } else {
      g.pdst[slot] = kEmpty;
    }
    if (lane == 0) {
      g.pcnt[leaf] = static_cast<uint32_t>(c);
      if (c == 0) {
        // empty leaf: fence just above the window's last key
        uint64_t m = tot - 1;
        uint64_t i = merge_split(AA, na, BB, nb, m);
        uint64_t jb = m - i;
        uint64_t e = (i < na && (jb >= nb || A[i] < B[jb])) ? A[i] : B[jb];
        g.fence[leaf] = e + 1;
      }
    }
  }
}

// Count tree rebuild, bottom-up from level L0: each block folds 512
// level-L0 nodes through up to 9 levels in LDS.  Replaces per-window /
// per-key atomics on shared ancestors.
__global__ void __launch_bounds__(256) k_tree_up(uint32_t *pcnt, uint32_t H, uint32_t L0) {
  __shared__ uint32_t sm[256];
  const uint64_t nin = 1ull << (H - L0);
  const uint64_t base_in = static_cast<uint64_t>(blockIdx.x) * 512;
  const uint32_t t = threadIdx.x;
  const uint64_t i0 = base_in + 2 * t;
  uint32_t v = 0;
  if (i0 < nin) v = pcnt[tree_level_off(H, L0) + i0] + (i0 + 1 < nin ? pcnt[tree_level_off(H, L0) + i0 + 1] : 0u);
  uint32_t lvl = L0 + 1;
  {
    uint64_t idx = (base_in >> 1) + t;
    if (idx < (1ull << (H - lvl))) pcnt[tree_level_off(H, lvl) + idx] = v;
  }
  sm[t] = v;
  __syncthreads();
  uint32_t width = 256;
  while (width > 1 && lvl < H) {
    width >>= 1;
    lvl++;
    uint32_t nv = 0;
    if (t < width) nv = sm[2 * t] + sm[2 * t + 1];
    __syncthreads();
    if (t < width) {
      sm[t] = nv;
      uint64_t idx = (base_in >> (lvl - L0)) + t;
      if (idx < (1ull << (H - lvl))) pcnt[tree_level_off(H, lvl) + idx] = nv;
    }
    __syncthreads();
  }
}

// Begin slot of vertex s from slot p on: the first occupied slot at or
// after p if it holds an element of s, else kPmaNone (the PMA is sorted, so
// the vertex has no element later).  A leaf at a time: its count and all
// 64 destinations loaded together, the first occupied slot found in
// registers.
__device__ inline uint64_t next_begin(const GView &g, uint32_t s, uint64_t p) {
  while (p < g.N) {
    const uint64_t leaf = p >> kLeafShift, base = leaf << kLeafShift;
    const uint32_t cnt = g.pcnt[leaf];
    const uint4 *pd = reinterpret_cast<const uint4 *>(g.pdst + base);
    uint4 v[kLeafSlots / 4];
#pragma unroll
    for (uint32_t j = 0; j < kLeafSlots / 4; j++) v[j] = pd[j];
    uint64_t found = ~0ull;
    if (cnt != 0) {
      const uint32_t from = static_cast<uint32_t>(p - base);
#pragma unroll
      for (uint32_t j = kLeafSlots / 4; j-- > 0;) {  // lowest index wins
        if (4 * j + 3 >= from && v[j].w != kEmpty) found = base + 4 * j + 3;
        if (4 * j + 2 >= from && v[j].z != kEmpty) found = base + 4 * j + 2;
        if (4 * j + 1 >= from && v[j].y != kEmpty) found = base + 4 * j + 1;
        if (4 * j >= from && v[j].x != kEmpty) found = base + 4 * j;
      }
    }
    if (found != ~0ull) return g.psrc[found] == s ? found : kPmaNone;
    p = base + kLeafSlots;
  }
  return kPmaNone;
}

// First slot at or after p holding an element of s that this batch does
// not delete, or kPmaNone.  keys[j, k) are the batch's PMA deletes after
// the caller's (sorted, so those of s come first); the slots they empty may
// read either way while the other threads work, so each element found is
// checked against them (a merge of two sorted lists).  Empty leaves are
// skipped by their counts.
__device__ inline uint64_t next_begin_after(const GView &g, uint32_t s, uint64_t p, const uint64_t *keys, uint64_t j,
                                            uint64_t k) {
  while (p < g.N) {
    const uint64_t leaf = p >> kLeafShift, end = (leaf + 1) << kLeafShift;
    if (g.pcnt[leaf] == 0) {
      p = end;
      continue;
    }
    for (; p < end; p++) {
      const uint32_t d = g.pdst[p];
      if (d == kEmpty) continue;
      if (g.psrc[p] != s) return kPmaNone;  // the PMA is sorted: no later element of s
      const uint64_t key = make_key(s, d);
      while (j < k && keys[j] < key) j++;
      if (j < k && keys[j] == key) continue;  // deleted by this batch
      return p;
    }
  }
  return kPmaNone;
}

// PMA delete: clear the slot and lower the leaf count.  The thread whose key
// is its vertex's first element also moves the vertex's begin pointer to
// the first element the batch keeps.  A leaf left at or below the leaf
// lower bound lo0 raises *low (the batch then respreads the underfull
// windows, as src/PMA.cc:1636-1746 does per removal).
// hint (optional): per key the slot where the batch's preparation found it.
// The insert half may have moved it since, so a hint is used only while the
// slot still holds the key; otherwise the key is searched for.
__global__ void k_pma_delete(GView g, const uint64_t *keys, uint64_t cap, uint64_t *err, const uint64_t *kdev,
                             uint32_t lo0, uint64_t *low, const uint64_t *hint) {
  const uint64_t k = dev_count(kdev, cap);
  GRID_STRIDE(i, k) {
    const uint32_t s = key_src(keys[i]), d = key_dst(keys[i]);
    const uint64_t begin = g.vb[s].aux;
    uint64_t slot = ~0ull;
    if (hint) {
      const uint64_t h = hint[i];
      if (h < g.N && g.pdst[h] == d && g.psrc[h] == s) slot = h;
    }
    if (slot == ~0ull) slot = pma_slot_of(g, s, d, begin);
    if (slot == ~0ull) {
      atomicOr(reinterpret_cast<unsigned long long *>(err), 4ull);
      continue;
    }
    g.pdst[slot] = kEmpty;
    const uint64_t leaf = slot >> kLeafShift;
    if (atomicSub(&g.pcnt[leaf], 1u) - 1u <= lo0) atomicOr(reinterpret_cast<unsigned long long *>(low), 1ull);
    if (begin == slot) g.vb[s].aux = next_begin_after(g, s, slot + 1, keys, i + 1, k);
  }
}

// ------------------------------------------------------------ promotion
// One wave per promoted vertex: copy its PMA run (D_old - 13 elements
// from the begin slot) into ext[ext_off ..] and remove it from the PMA
// (slots emptied, leaf counts lowered, *low raised when a leaf falls to
// lo0 or below).  The vertex's begin word is replaced by its tree next.
__global__ void __launch_bounds__(256) k_extract_pma(GView g, const uint32_t *pv, const uint32_t *pcount,
                                                     const uint32_t *ext_off, uint32_t np, uint64_t *ext,
                                                     uint32_t lo0, uint64_t *low) {
  uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  uint32_t lane = threadIdx.x & 63;
  for (uint64_t w = wave; w < np; w += nwaves) {
    uint32_t s = pv[w];
    uint32_t need = pcount[w];
    if (need == 0) continue;
    uint64_t p = g.vb[s].aux;
    uint32_t got = 0;
    while (got < need) {
      uint64_t slot = p + lane;
      uint32_t d = slot < g.N ? g.pdst[slot] : kEmpty;
      bool v = d != kEmpty;
      uint64_t m = __ballot(v);
      uint32_t r = __popcll(m & lanemask_lt());
      if (v && got + r < need) {
        ext[ext_off[w] + got + r] = make_key(s, d);
        g.pdst[slot] = kEmpty;
        if (atomicSub(&g.pcnt[slot >> kLeafShift], 1u) - 1u <= lo0)
          atomicOr(reinterpret_cast<unsigned long long *>(low), 1ull);
      }
      got += __popcll(m);
      p += 64;
      if (p >= g.N) break;
    }
  }
}

__global__ void k_promote_info(const uint64_t *R, const uint32_t *run_beg, const uint32_t *promote,
                               const uint32_t *ppos, uint32_t nruns, GView g, uint32_t *pv, uint32_t *pcount) {
  GRID_STRIDE(r, nruns) if (promote[r]) {
    uint32_t s = key_src(R[run_beg[r]]);
    pv[ppos[r]] = s;
    // degree was already raised by this batch's new keys; the PMA run
    // holds the old secondary count (old degree - 13).
    pcount[ppos[r]] = 0;
  }
}

__global__ void k_promote_old_counts(const uint32_t *pv, uint32_t np, const uint32_t *old_deg, uint32_t *pcount) {
  GRID_STRIDE(i, np) {
    uint32_t D = old_deg[i];
    pcount[i] = D > kInline ? D - kInline : 0;
  }
}

__global__ void k_set_tree(GView g, const uint32_t *pv, uint32_t np, uint32_t first_tree, uint32_t *tvtx) {
  GRID_STRIDE(i, np) {
    uint32_t t = first_tree + static_cast<uint32_t>(i);
    g.vb[pv[i]].aux = kAuxTree | t;
    g.tnblk[t] = 0;
    g.toff[t] = 0;
    tvtx[t] = pv[i];
  }
}

// ------------------------------------------------------------ tree kernels
__global__ void k_tree_loc(GView g, const uint64_t *keys, uint64_t k, uint64_t *gj) {
  GRID_STRIDE(i, k) {
    uint32_t s = key_src(keys[i]), d = key_dst(keys[i]);
    uint32_t t = static_cast<uint32_t>(g.vb[s].aux);
    uint32_t j = g.tnblk[t] == 0 ? 0 : tree_block_pos(g, t, d);
    gj[i] = (static_cast<uint64_t>(t) << 32) | j;
  }
}

__global__ void k_flag_change_u64(const uint64_t *a, uint64_t n, uint32_t *flags) {
  GRID_STRIDE(i, n) flags[i] = (i == 0 || a[i] != a[i - 1]) ? 1u : 0u;
}

struct Grp {
  uint32_t t, j;
  uint32_t kb, ke;
  uint32_t old_b;  // ~0 when the tree is empty
  uint32_t ocnt;
  uint32_t ob;     // output blocks
  uint32_t has_old;
};

__global__ void k_make_groups(GView g, const uint64_t *gj, const uint32_t *flags, const uint32_t *pos, uint64_t k,
                              uint32_t ng, Grp *grp, uint32_t *grp_kb) {
  GRID_STRIDE(i, k) if (flags[i]) {
    uint32_t x = pos[i];
    grp_kb[x] = static_cast<uint32_t>(i);
    Grp G;
    G.t = static_cast<uint32_t>(gj[i] >> 32);
    G.j = static_cast<uint32_t>(gj[i]);
    G.kb = static_cast<uint32_t>(i);
    grp[x] = G;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) grp_kb[ng] = static_cast<uint32_t>(k);
}

// insert=true: total = old + new; delete: total = old - deleted.
__global__ void k_group_sizes(GView g, Grp *grp, const uint32_t *grp_kb, uint32_t ng, bool insert, uint32_t *ob,
                              uint32_t *has_old) {
  GRID_STRIDE(x, ng) {
    Grp G = grp[x];
    G.ke = grp_kb[x + 1];
    uint32_t nb = g.tnblk[G.t];
    G.has_old = nb > 0 ? 1u : 0u;
    G.old_b = nb > 0 ? g.iblk[g.toff[G.t] + G.j] : ~0u;
    G.ocnt = nb > 0 ? g.bcnt[G.old_b] : 0u;
    uint32_t kcount = G.ke - G.kb;
    uint32_t total = insert ? G.ocnt + kcount : G.ocnt - kcount;
    if (insert) G.ob = total <= kBlk ? 1u : (total + kBlkFill - 1) / kBlkFill;
    else G.ob = total > 0 ? 1u : 0u;
    grp[x] = G;
    ob[x] = G.ob;
    has_old[x] = G.has_old;
  }
}

__device__ inline uint32_t find_group_of_block(const uint32_t *obpre, uint32_t ng, uint64_t x) {
  uint32_t lo = 0, hi = ng;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (obpre[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

// One 256-thread block per output leaf block of a tree insert.
__global__ void __launch_bounds__(256) k_tree_fill_insert(GView g, const Grp *grp, const uint32_t *obpre,
                                                          uint32_t ng, uint64_t nout, uint64_t blk_base,
                                                          const uint64_t *keys) {
  for (uint64_t x = blockIdx.x; x < nout; x += gridDim.x) {
    // groups with ob == 0 never occur on insert, so the search is exact
    uint32_t gi = find_group_of_block(obpre, ng, x);
    Grp G = grp[gi];
    uint64_t bi = x - obpre[gi];
    uint64_t na = G.ocnt, nb = G.ke - G.kb;
    uint64_t tot = na + nb;
    uint64_t q = tot / G.ob, rem = tot % G.ob;
    uint64_t c = q + (bi < rem ? 1 : 0);
    uint64_t m0 = bi * q + (bi < rem ? bi : rem);
    const uint32_t *A = G.has_old ? g.bkeys + static_cast<uint64_t>(G.old_b) * kBlk : nullptr;
    U32Arr AA{A};
    DstArr BB{keys + G.kb};
    uint64_t nbk = blk_base + x;
    for (uint64_t e = threadIdx.x; e < c; e += blockDim.x) {
      uint64_t m = m0 + e;
      uint64_t i = merge_split(AA, na, BB, nb, m);
      uint64_t jb = m - i;
      uint32_t v = (i < na && (jb >= nb || A[i] < BB(jb))) ? A[i] : BB(jb);
      g.bkeys[nbk * kBlk + e] = v;
    }
    if (threadIdx.x == 0) g.bcnt[nbk] = static_cast<uint32_t>(c);
  }
}

// One 256-thread block per group of a tree delete (groups with ob == 1).
__global__ void __launch_bounds__(256) k_tree_fill_delete(GView g, const Grp *grp, const uint32_t *obpre,
                                                          uint32_t ng, uint64_t blk_base, const uint64_t *keys) {
  __shared__ uint32_t wsum[4];
  for (uint64_t x = blockIdx.x; x < ng; x += gridDim.x) {
    Grp G = grp[x];
    if (G.ob == 0) continue;
    const uint32_t *A = g.bkeys + static_cast<uint64_t>(G.old_b) * kBlk;
    uint64_t nbk = blk_base + obpre[x];
    uint32_t e = threadIdx.x;
    bool keep = false;
    uint32_t v = 0;
    if (e < G.ocnt) {
      v = A[e];
      uint32_t lo = G.kb, hi = G.ke;
      while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (key_dst(keys[mid]) < v) lo = mid + 1;
        else hi = mid;
      }
      keep = !(lo < G.ke && key_dst(keys[lo]) == v);
    }
    uint64_t m = __ballot(keep);
    uint32_t r = __popcll(m & lanemask_lt());
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) base += wsum[w];
    if (keep) g.bkeys[nbk * kBlk + base + r] = v;
    if (threadIdx.x == 0) g.bcnt[nbk] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
}

struct TreeRun {
  uint32_t t;
  uint32_t g0, g1;
  uint32_t nb_old;
  uint64_t new_off;
  uint32_t nb_new;
};

__global__ void k_tree_runs(GView g, const Grp *grp, const uint32_t *flags, const uint32_t *pos, uint32_t ng,
                            uint32_t ntr, TreeRun *tr, uint32_t *tr_start) {
  GRID_STRIDE(x, ng) if (flags[x]) {
    tr_start[pos[x]] = static_cast<uint32_t>(x);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) tr_start[ntr] = ng;
}

__global__ void k_tree_run_sizes(GView g, const Grp *grp, const uint32_t *tr_start, uint32_t ntr,
                                 const uint32_t *obpre, const uint32_t *hopre, TreeRun *tr, uint32_t *nb_new,
                                 uint32_t *nb_old) {
  GRID_STRIDE(r, ntr) {
    TreeRun R;
    R.g0 = tr_start[r];
    R.g1 = tr_start[r + 1];
    R.t = grp[R.g0].t;
    R.nb_old = g.tnblk[R.t];
    int64_t delta = static_cast<int64_t>(obpre[R.g1] - obpre[R.g0]) - static_cast<int64_t>(hopre[R.g1] - hopre[R.g0]);
    R.nb_new = static_cast<uint32_t>(static_cast<int64_t>(R.nb_old) + delta);
    R.new_off = 0;
    tr[r] = R;
    nb_new[r] = R.nb_new;
    nb_old[r] = R.nb_old;
  }
}

__device__ inline uint32_t find_run(const uint32_t *pre, uint32_t n, uint64_t j) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= j) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

// Copy the untouched entries of each touched tree's index into its new
// segment, shifted by the blocks that earlier groups added or removed.
__global__ void k_tree_idx_copy(GView g, const TreeRun *tr, const uint32_t *oldpre, uint32_t ntr, uint64_t ntask,
                                const Grp *grp, const uint32_t *obpre, const uint32_t *hopre,
                                const uint32_t *newpre, uint64_t idx_base) {
  GRID_STRIDE(j, ntask) {
    uint32_t r = find_run(oldpre, ntr, j);
    TreeRun R = tr[r];
    uint32_t e = static_cast<uint32_t>(j - oldpre[r]);
    // first group of this tree with grp.j >= e
    uint32_t lo = R.g0, hi = R.g1;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (grp[mid].j < e) lo = mid + 1;
      else hi = mid;
    }
    if (lo < R.g1 && grp[lo].j == e) continue;  // rewritten by its group
    int64_t shift = static_cast<int64_t>(obpre[lo] - obpre[R.g0]) - static_cast<int64_t>(hopre[lo] - hopre[R.g0]);
    uint64_t dst = idx_base + newpre[r] + static_cast<uint64_t>(static_cast<int64_t>(e) + shift);
    uint64_t src = g.toff[R.t] + e;
    g.iblk[dst] = g.iblk[src];
    g.imin[dst] = g.imin[src];
  }
}

__global__ void k_tree_idx_groups(GView g, const TreeRun *tr, uint32_t ntr, const Grp *grp,
                                  const uint32_t *grp_run, uint32_t ng, const uint32_t *obpre,
                                  const uint32_t *hopre, const uint32_t *newpre, uint64_t idx_base,
                                  uint64_t blk_base) {
  GRID_STRIDE(x, ng) {
    Grp G = grp[x];
    TreeRun R = tr[grp_run[x]];
    int64_t shift = static_cast<int64_t>(obpre[x] - obpre[R.g0]) - static_cast<int64_t>(hopre[x] - hopre[R.g0]);
    for (uint32_t b = 0; b < G.ob; b++) {
      uint64_t nbk = blk_base + obpre[x] + b;
      uint64_t dst = idx_base + newpre[grp_run[x]] + static_cast<uint64_t>(static_cast<int64_t>(G.j) + shift + b);
      g.iblk[dst] = static_cast<uint32_t>(nbk);
      g.imin[dst] = g.bkeys[nbk * kBlk];
    }
  }
}

__global__ void k_grp_run(const uint32_t *flags, const uint32_t *pos, uint32_t ng, uint32_t *grp_run) {
  GRID_STRIDE(x, ng) grp_run[x] = pos[x] + flags[x] - 1;
}

__global__ void k_tree_commit(GView g, const TreeRun *tr, uint32_t ntr, const uint32_t *newpre, uint64_t idx_base) {
  GRID_STRIDE(r, ntr) {
    TreeRun R = tr[r];
    g.toff[R.t] = idx_base + newpre[r];
    g.tnblk[R.t] = R.nb_new;
  }
}

// ------------------------------------------------------------ small tree updates in place
// A tree update of at most Graph::kTreeSmallKeys keys whose groups (the keys that
// land in one leaf block) all fit their block -- an insert leaves it with at
// most kBlk keys, a delete leaves at least one -- rewrites those blocks in
// place: two launches and one read instead of the copy-on-write pass, whose
// block and index counts need three reads.  The index keeps its blocks; only
// a block's count and first key change.  When a group does not fit, nothing
// is written and the copy-on-write pass runs.
// One workgroup.  gl[0] = number of groups (0 when one does not fit), then
// per group x: gl[1 + 3x] first key, gl[2 + 3x] end, gl[3 + 3x] index
// position of its block (groups in any order).  *status = 1 when a group
// does not fit, else 0.
template <bool kInsert>
__global__ void __launch_bounds__(1024) k_tree_small_groups(GView g, const uint64_t *keys, uint32_t cap, uint32_t *gl,
                                                            uint64_t *status, const uint64_t *kdev,
                                                            const uint64_t *gate) {
  __shared__ uint64_t s_gj[Graph::kTreeSmallKeys];
  __shared__ uint32_t s_ng, s_fail;
  // kdev: the key count computed on the device (null: cap); more than
  // kTreeSmallKeys keys do not fit (status 1).  gate: when *gate is nonzero
  // the update must wait for the host (status 2); nothing is written.
  const uint64_t kk = kdev ? *kdev : cap;
  const uint32_t refuse = gate && *gate ? 2u : (kk > Graph::kTreeSmallKeys || kk > cap ? 1u : 0u);
  if (refuse) {
    if (threadIdx.x == 0) {
      gl[0] = 0;
      *status = refuse;
    }
    return;
  }
  const uint32_t k = static_cast<uint32_t>(kk);
  if (threadIdx.x == 0) {
    s_ng = 0;
    s_fail = 0;
  }
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    const uint64_t key = keys[i];
    const uint32_t t = static_cast<uint32_t>(g.vb[key_src(key)].aux);
    const uint32_t j = g.tnblk[t] == 0 ? 0 : tree_block_pos(g, t, key_dst(key));
    s_gj[i] = (static_cast<uint64_t>(t) << 32) | j;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    if (i > 0 && s_gj[i] == s_gj[i - 1]) continue;
    // a group longer than kBlk + 1 keys never fits
    uint32_t e = i + 1;
    while (e < k && e - i <= kBlk && s_gj[e] == s_gj[i]) e++;
    const uint32_t t = static_cast<uint32_t>(s_gj[i] >> 32), j = static_cast<uint32_t>(s_gj[i]);
    bool ok = g.tnblk[t] != 0 && e - i <= kBlk;
    if (ok) {
      const uint32_t c = g.bcnt[g.iblk[g.toff[t] + j]];
      ok = kInsert ? c + (e - i) <= kBlk : c > e - i;
    }
    if (!ok) {
      s_fail = 1;
    } else {
      const uint32_t x = atomicAdd(&s_ng, 1u);
      gl[1 + 3 * x] = i;
      gl[2 + 3 * x] = e;
      gl[3 + 3 * x] = j;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    gl[0] = s_fail ? 0 : s_ng;
    *status = s_fail;
  }
}

// One 256-thread workgroup per group listed by k_tree_small_groups: the
// group's block rewritten in place (old keys staged in LDS first).
template <bool kInsert>
__global__ void __launch_bounds__(256) k_tree_small_apply(GView g, const uint64_t *keys, const uint32_t *gl) {
  __shared__ uint32_t s_old[kBlk];
  __shared__ uint32_t s_wsum[4];
  const uint32_t ng = gl[0];
  for (uint32_t x = blockIdx.x; x < ng; x += gridDim.x) {
    const uint32_t i0 = gl[1 + 3 * x], i1 = gl[2 + 3 * x], j = gl[3 + 3 * x];
    const uint32_t t = static_cast<uint32_t>(g.vb[key_src(keys[i0])].aux);
    const uint64_t io = g.toff[t] + j;
    const uint32_t b = g.iblk[io];
    const uint32_t c = g.bcnt[b];
    const uint32_t cnt = i1 - i0;
    uint32_t *blk = g.bkeys + static_cast<uint64_t>(b) * kBlk;
    const uint32_t p = threadIdx.x;
    if (p < c) s_old[p] = blk[p];
    __syncthreads();
    // position of v among the group's keys (sorted by destination)
    auto rank_in_keys = [&](uint32_t v) {
      uint32_t lo = i0, hi = i1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (key_dst(keys[mid]) < v) lo = mid + 1;
        else hi = mid;
      }
      return lo;
    };
    if (kInsert) {
      // merged order: an old key moves right by the new keys below it, a
      // new key lands after the old keys below it (the sets are disjoint)
      if (p < c) {
        const uint32_t v = s_old[p];
        blk[p + (rank_in_keys(v) - i0)] = v;
      }
      for (uint32_t q = p; q < cnt; q += blockDim.x) {
        const uint32_t v = key_dst(keys[i0 + q]);
        uint32_t lo = 0, hi = c;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_old[mid] < v) lo = mid + 1;
          else hi = mid;
        }
        blk[q + lo] = v;
      }
      if (p == 0) {
        g.bcnt[b] = c + cnt;
        const uint32_t k0 = key_dst(keys[i0]);
        if (k0 < s_old[0]) g.imin[io] = k0;
      }
    } else {
      bool keep = false;
      uint32_t v = 0;
      if (p < c) {
        v = s_old[p];
        const uint32_t r = rank_in_keys(v);
        keep = !(r < i1 && key_dst(keys[r]) == v);
      }
      const uint64_t m = __ballot(keep);
      const uint32_t r = __popcll(m & lanemask_lt());
      if ((p & 63) == 0) s_wsum[p >> 6] = __popcll(m);
      __syncthreads();
      uint32_t base = 0;
      for (uint32_t w = 0; w < (p >> 6); w++) base += s_wsum[w];
      if (keep) {
        blk[base + r] = v;
        if (base + r == 0) g.imin[io] = v;
      }
      if (p == 0) g.bcnt[b] = s_wsum[0] + s_wsum[1] + s_wsum[2] + s_wsum[3];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ delete batch
// The delete half's first kernel, one thread per run of deleted keys (exf:
// their existence flags and dnidx the prefix of the absent ones; null when
// every key of R exists): its
// gate (gate null: open; spec: also the insert half of the same batch must
// have completed) and counts (cw, published at gw); the run's inline hits
// removed (kept entries move down over them, the degree drops by the run's
// length); how many inline slots the refill must fill (run_bump); and the
// run's secondary deletes written, in key order, to its offset in the PMA
// list P or the B-tree list T.  The offsets come from the run scan, as do
// each B-tree run's offset for the tree keys its refill moves (tbump_off).
// Totals: ds[4] PMA deletes, ds[5] tree deletes, ds[6] inline refills,
// ds[17] refill keys for the tree delete (ds[16], refill keys for the PMA,
// is 0: k_bump removes those itself); ds[3] the existing keys deleted.
// ds[9] (k_bump's counter of PMA elements moved inline) and, when the half
// runs, *low are cleared here.  dslot (optional): per key of R its PMA slot
// as the preparation found it; each key written to P takes its slot to Ph.
__global__ void __launch_bounds__(kRunBlock) k_delete_runs(GView g, const uint64_t *R, const uint32_t *exf,
                                                           const uint32_t *dnidx, const uint32_t *run_beg,
                                                           const uint64_t *cw, const uint64_t *gate, bool spec,
                                                           bool pma_launched, uint64_t *gw, uint64_t cap_keys,
                                                           uint32_t cap_runs, uint64_t *P, uint64_t *T,
                                                           uint32_t *run_bump, uint32_t *tbump_off, uint64_t *ds,
                                                           uint64_t *low, const uint64_t *dslot, uint64_t *Ph,
                                                           uint64_t *state, uint32_t *misc,
                                                           uint32_t *err, uint32_t nblocks, uint32_t epoch) {
  __shared__ uint32_t s_ticket;
  __shared__ uint32_t s_b0[kRunBlock + 1], s_off[kRunBlock], s_nb0[kRunBlock];
  __shared__ uint8_t s_tier[kRunBlock], s_nhit[kRunBlock];
  __shared__ uint32_t s_hit[kRunBlock][kInline];  // the inline entries the run deletes, sorted
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  const bool ok = !gate || (prep_clean(gate) && (!spec || insert_half_done(gate, pma_launched)));
  const HalfCounts hc = gate_counts(ok, cw, cap_keys, cap_runs, gw);
  const uint32_t t = threadIdx.x;
  if (blockIdx.x == 0 && t == 0) {
    ds[9] = 0;
    ds[3] = exf ? hc.keys - dnidx[hc.keys] : hc.keys;  // the existing keys deleted
    if (ok) *low = 0;
  }
  const uint64_t r0 = static_cast<uint64_t>(b) * kRunBlock;
  const uint32_t nlive = r0 < hc.runs ? static_cast<uint32_t>(hc.runs - r0 < kRunBlock ? hc.runs - r0 : kRunBlock) : 0u;
  const bool live = t < nlive;
  const uint64_t r = r0 + t;
  uint32_t nsec = 0, y = 0, nhit = 0, nb0 = 0;
  uint8_t tier = kRoutePma;
  if (live) {
    const uint32_t b0 = run_beg[r], b1 = run_beg[r + 1];
    nb0 = exf ? dnidx[b0] : 0u;
    // existing keys of the run (an absent key matches no inline entry)
    const uint32_t len = (b1 - b0) - (exf ? dnidx[b1] - nb0 : 0u);
    const uint32_t s = key_src(R[b0]);
    // the block as 16 words (degree, nbr[0..12], aux): registers only, the
    // inline hits as a 13-bit mask (no run-time indexed private array)
    const uint4 *vp = reinterpret_cast<const uint4 *>(g.vb + s);
    const uint4 q0 = vp[0], q1 = vp[1], q2 = vp[2], q3 = vp[3];
    const uint32_t w[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                            q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
    const uint32_t D = w[0];
    const uint32_t lim = D < kInline ? D : kInline;
    tier = (w[15] & 0x80000000u) ? kRouteTree : kRoutePma;  // aux bit 63: B-tree
    uint32_t hit = 0;
    if (b1 - b0 <= 2 * kInline) {
      for (uint32_t q = b0; q < b1; q++) {
        const uint32_t d = key_dst(R[q]);
#pragma unroll
        for (uint32_t p = 0; p < kInline; p++) hit |= (p < lim && w[1 + p] == d) ? (1u << p) : 0u;
      }
    } else {
      // a long run: look each inline entry up among the run's keys
#pragma unroll
      for (uint32_t p = 0; p < kInline; p++) {
        if (p >= lim) continue;
        const uint64_t key = make_key(s, w[1 + p]);
        const uint64_t lb = b0 + lower_bound_dev(R + b0, b1 - b0, key);
        if (lb < b1 && R[lb] == key) hit |= 1u << p;
      }
    }
    // kept inline entries move down over the deleted ones, in order
    VBlock &vb = g.vb[s];
    uint32_t keep = 0;
#pragma unroll
    for (uint32_t p = 0; p < kInline; p++) {
      if (p < lim && !((hit >> p) & 1u)) {
        if (keep != p) vb.nbr[keep] = w[1 + p];
        keep++;
      } else if (p < lim) {
        s_hit[t][nhit++] = w[1 + p];
      }
    }
    const uint32_t x = lim - keep;  // inline entries deleted
    const uint32_t sec_before = D > kInline ? D - kInline : 0;
    const uint32_t sec_after = sec_before - (len - x);
    y = x < sec_after ? x : sec_after;
    if (len) vb.degree = D - len;
    run_bump[r] = y;
    nsec = len - x;
    s_b0[t] = b0;
    if (t == 0) s_b0[nlive] = run_beg[r0 + nlive];  // run_beg[nruns] = nR
  }
  const uint32_t xs[4] = {tier == kRoutePma ? nsec : 0u, tier == kRouteTree ? nsec : 0u, y,
                          tier == kRouteTree ? y : 0u};
  uint32_t pre[4], tot[4];
  run_block_scan<4>(xs, pre, tot, state, b, err, epoch);
  if (live) tbump_off[r] = pre[3];
  s_nb0[t] = nb0;
  s_off[t] = tier == kRoutePma ? pre[0] : pre[1];
  s_tier[t] = tier;
  s_nhit[t] = static_cast<uint8_t>(nhit);
  __syncthreads();
  if (b == nblocks - 1 && t == 0) {
    ds[4] = tot[0];
    ds[5] = tot[1];
    ds[6] = tot[2];
    ds[16] = 0;
    ds[17] = tot[3];
  }
  if (nlive == 0) return;
  // the secondary deletes, by every thread over the block's keys: a key's
  // position is its rank among the run's existing keys minus the inline
  // hits below it
  const uint32_t kb = s_b0[0], ke = s_b0[nlive];
  for (uint32_t i = kb + t; i < ke; i += kRunBlock) {
    if (exf && !exf[i]) continue;
    uint32_t lo = 0, hi = nlive;  // run: last q with s_b0[q] <= i
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_b0[mid] <= i) lo = mid;
      else hi = mid;
    }
    const uint32_t d = key_dst(R[i]);
    const uint32_t nh = s_nhit[lo];
    uint32_t less = 0;
    bool inl = false;
    for (uint32_t e = 0; e < nh; e++) {
      const uint32_t v = s_hit[lo][e];
      less += v < d ? 1u : 0u;
      inl |= v == d;
    }
    if (inl) continue;
    const uint32_t rank = (i - s_b0[lo]) - (exf ? dnidx[i] - s_nb0[lo] : 0u);
    const uint32_t pos = s_off[lo] + rank - less;
    if (s_tier[lo] == kRoutePma) {
      P[pos] = R[i];
      if (dslot) Ph[pos] = dslot[i];
    } else {
      T[pos] = R[i];
    }
  }
}

// Refill inline slots from the smallest secondary neighbours (graph.h:729-780,
// batched): thread per run with y > 0.
// PMA vertices: the moved elements are removed from the PMA here (slot
// emptied, leaf count lowered, begin pointer moved to the next element of
// the vertex; their count added to *pma_taken).
// B-tree vertices: the moved keys are written to tree_keys at the run's
// offset (tbump_off) for the tree delete.  With trees_only, PMA vertices
// are skipped (a second pass after the B-tree fallback; their refill is
// already done).  gw: the counts k_delete_runs published.
__global__ void k_bump(GView g, const uint64_t *R, const uint32_t *run_beg, const uint64_t *gw,
                       const uint32_t *run_bump, const uint32_t *tbump_off, uint64_t *tree_keys, bool trees_only,
                       uint64_t *pma_taken, uint32_t lo0, uint64_t *low) {
  const uint32_t nruns = static_cast<uint32_t>(gw[1]);
  GRID_STRIDE(r, nruns) {
    uint32_t y = run_bump[r];
    if (y == 0) continue;
    uint32_t s = key_src(R[run_beg[r]]);
    VBlock &vb = g.vb[s];
    if (trees_only && !(vb.aux & kAuxTree)) continue;
    uint32_t D = vb.degree;  // already the post-delete degree
    // inline currently holds min(D,13) - y entries
    uint32_t have = (D < kInline ? D : kInline) - y;
    uint32_t got = 0;
    if (vb.aux & kAuxTree) {
      uint32_t t = static_cast<uint32_t>(vb.aux);
      uint64_t off = g.toff[t];
      uint32_t nb = g.tnblk[t];
      uint64_t out = tbump_off[r];
      for (uint32_t j = 0; j < nb && got < y; j++) {
        uint32_t b = g.iblk[off + j];
        uint32_t c = g.bcnt[b];
        for (uint32_t e = 0; e < c && got < y; e++) {
          uint32_t d = g.bkeys[static_cast<uint64_t>(b) * kBlk + e];
          vb.nbr[have + got] = d;
          tree_keys[out + got] = make_key(s, d);
          got++;
        }
      }
    } else {
      uint64_t p = vb.aux;
      while (got < y && p < g.N) {
        uint32_t d = g.pdst[p];
        if (d != kEmpty) {
          vb.nbr[have + got] = d;
          g.pdst[p] = kEmpty;
          if (atomicSub(&g.pcnt[p >> kLeafShift], 1u) - 1u <= lo0)
            atomicOr(reinterpret_cast<unsigned long long *>(low), 1ull);
          got++;
        }
        p++;
      }
      vb.aux = next_begin(g, s, p);  // elements after the vertex's belong to later vertices
      atomicAdd(reinterpret_cast<unsigned long long *>(pma_taken), static_cast<unsigned long long>(got));
    }
  }
}

// ------------------------------------------------------------ count_common
// Graph::count_common (graph.h:901-1255): |N(s) ∩ N(d) ∩ [0, min(s, d))|,
// the common neighbours below both endpoints, over the six tier pairs.  One
// wave per pair: it walks the sorted list of the endpoint with the smaller
// degree, 64 neighbours per step (inline slots, then its PMA run or its tree
// blocks), up to the first neighbour >= min(s, d); each lane tests its
// neighbour against the other endpoint with the is_edge lookup.  Every lane
// returns the pair's count.
__device__ inline uint32_t count_common_wave(const GView &g, uint32_t s, uint32_t d) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t lim = s < d ? s : d;
  const uint32_t a = g.vb[s].degree <= g.vb[d].degree ? s : d;
  const uint32_t b = a == s ? d : s;
  const VBlock &blk = g.vb[a];
  const uint32_t D = blk.degree;
  uint32_t c = 0;
  if (lane < (D < kInline ? D : kInline)) {
    const uint32_t w = blk.nbr[lane];
    if (w < lim && edge_exists(g, b, w)) c++;
  }
  if (D > kInline && blk.nbr[kInline - 1] < lim) {
    const uint64_t aux = blk.aux;
    if (aux & kAuxTree) {
      const uint32_t t = static_cast<uint32_t>(aux);
      const uint64_t io = g.toff[t];
      const uint32_t nb = g.tnblk[t];
      bool more = true;
      for (uint32_t j = 0; j < nb && more; j++) {
        const uint32_t bid = g.iblk[io + j];
        const uint32_t cnt = g.bcnt[bid];
        for (uint32_t e = 0; e < cnt; e += 64) {
          const uint32_t idx = e + lane;
          const uint32_t w = idx < cnt ? g.bkeys[static_cast<uint64_t>(bid) * kBlk + idx] : kEmpty;
          if (w < lim && edge_exists(g, b, w)) c++;
          if (__ballot(idx < cnt && w >= lim)) {
            more = false;
            break;
          }
        }
      }
    } else {
      const uint32_t need = D - kInline;
      uint64_t p = aux;
      uint32_t got = 0;
      while (got < need && p < g.N) {
        const uint64_t slot = p + lane;
        const uint32_t w = slot < g.N ? g.pdst[slot] : kEmpty;
        const bool ok = w != kEmpty;
        const uint64_t m = __ballot(ok);
        const uint32_t r = __popcll(m & lanemask_lt());
        const bool take = ok && got + r < need;
        if (take && w < lim && edge_exists(g, b, w)) c++;
        if (__ballot(take && w >= lim)) break;
        got += __popcll(m);
        p += 64;
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  return c;
}

// One wave per (s[i], d[i]); ids at or above n give 0 and set *err.
__global__ void __launch_bounds__(256) k_count_common(GView g, const uint32_t *src, const uint32_t *dst, uint64_t q,
                                                      uint32_t *out, uint64_t *err) {
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  const uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t i = wave; i < q; i += nwaves) {
    const uint32_t s = src[i], d = dst[i];
    if (s >= g.n || d >= g.n) {
      if ((threadIdx.x & 63) == 0) {
        out[i] = 0;
        atomicOr(reinterpret_cast<unsigned long long *>(err), 1ull);
      }
      continue;
    }
    const uint32_t c = count_common_wave(g, s, d);
    if ((threadIdx.x & 63) == 0) out[i] = c;
  }
}

// TC.h's edgeMap with countF (TC.h:51-87): over the frontier's vertices u
// (list, or every vertex when null) and their neighbours v < u, the sum of
// count_common(u, v).  Edges are listed first from a CSR export (pairs of
// the frontier vertex's edges below it), then one wave per pair, so a hub's
// edges spread over the whole GPU.
// lower[i]: neighbours of the i-th frontier vertex below it.
__global__ void k_tc_lower(const uint64_t *off, const uint32_t *nbr, const uint32_t *list, uint64_t cnt,
                          uint32_t *lower) {
  GRID_STRIDE(i, cnt) {
    const uint32_t u = list ? list[i] : static_cast<uint32_t>(i);
    lower[i] = static_cast<uint32_t>(lower_bound_dev<uint32_t>(nbr + off[u], off[u + 1] - off[u], u));
  }
}

// One wave per frontier vertex writes its pairs (u, v) for v < u.
__global__ void __launch_bounds__(256) k_tc_pairs(const uint64_t *off, const uint32_t *nbr, const uint32_t *list,
                                                  uint64_t cnt, const uint64_t *pos, uint64_t *pairs) {
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  const uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t i = wave; i < cnt; i += nwaves) {
    const uint32_t u = list ? list[i] : static_cast<uint32_t>(i);
    const uint64_t b = off[u], k = pos[i + 1] - pos[i];
    for (uint64_t e = lane; e < k; e += 64) pairs[pos[i] + e] = make_key(u, nbr[b + e]);
  }
}

// One wave per pair; per-block sums, then one 64-bit atomic per block.
__global__ void __launch_bounds__(256) k_tc_count_pairs(GView g, const uint64_t *pairs, uint64_t np,
                                                        unsigned long long *total) {
  __shared__ unsigned long long part[4];
  const uint64_t wave = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 6;
  const uint64_t nwaves = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 6;
  unsigned long long acc = 0;
  for (uint64_t i = wave; i < np; i += nwaves) {
    const uint64_t pr = pairs[i];
    acc += count_common_wave(g, key_src(pr), key_dst(pr));
  }
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(total, t);
  }
}

// ------------------------------------------------------------ queries
__global__ void k_degrees(GView g, uint32_t *out) {
  GRID_STRIDE(v, g.n) out[v] = g.vb[v].degree;
}
__global__ void k_degrees_u64(GView g, uint64_t *out) {
  GRID_STRIDE(v, g.n) out[v] = g.vb[v].degree;
}

__global__ void k_is_edge(GView g, const uint32_t *src, const uint32_t *dst, uint64_t k, uint8_t *out) {
  GRID_STRIDE(i, k) {
    uint32_t s = src[i], d = dst[i];
    out[i] = (s < g.n && d < g.n && edge_exists(g, s, d)) ? 1 : 0;
  }
}

__global__ void k_tiers(GView g, uint8_t *out) {
  GRID_STRIDE(v, g.n) {
    const VBlock &b = g.vb[v];
    out[v] = (b.aux & kAuxTree) ? 2 : (b.degree > kInline ? 1 : 0);
  }
}

// [0] PMA-tier vertices, [1] their secondary edges, [2] B-tree vertices,
// [3] their secondary edges (device counters, zeroed by the caller)
__global__ void k_tier_counts(GView g, unsigned long long *out) {
  unsigned long long c[4] = {0, 0, 0, 0};
  GRID_STRIDE(v, g.n) {
    const VBlock &b = g.vb[v];
    bool t = (b.aux & kAuxTree) != 0;
    uint32_t sec = b.degree > kInline ? b.degree - kInline : 0;
    if (t) {
      c[2]++;
      c[3] += sec;
    } else if (b.degree > kInline) {
      c[0]++;
      c[1] += sec;
    }
  }
  for (int j = 0; j < 4; j++) {
    unsigned long long x = c[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(out + j, x);
  }
}

// 16 lanes per vertex: lane i writes inline slot i (a group's stores are
// one run of consecutive ids).
__global__ void __launch_bounds__(256) k_export_inline(GView g, const uint64_t *off, uint32_t *nbrs) {
  const uint32_t gl = threadIdx.x & 15;
  const uint64_t grp = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 4;
  const uint64_t ngrp = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 4;
  for (uint64_t v = grp; v < g.n; v += ngrp) {
    const VBlock &b = g.vb[v];
    const uint32_t lim = b.degree < kInline ? b.degree : kInline;
    if (gl < lim) nbrs[off[v] + gl] = b.nbr[gl];
  }
}

// 16-lane group per PMA vertex, 4 slots per lane per 16-byte load (64 slots
// per group step, four vertices per wave); ballot ranks keep slot order.
__global__ void __launch_bounds__(256) k_export_pma(GView g, const uint32_t *list, uint32_t cnt, const uint64_t *off,
                                                    uint32_t *nbrs) {
  const uint32_t lane = threadIdx.x & 63, gl = lane & 15;
  const uint64_t gmask = 0xFFFFull << (lane & 48);
  const uint64_t lt = lanemask_lt() & gmask;
  const uint64_t grp = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) >> 4;
  const uint64_t ngrp = (static_cast<uint64_t>(gridDim.x) * blockDim.x) >> 4;
  for (uint64_t w = grp; w < cnt; w += ngrp) {
    const uint32_t v = list[w];
    const VBlock &b = g.vb[v];
    const uint32_t need = b.degree - kInline;
    const uint64_t begin = b.aux;
    const uint64_t o = off[v] + kInline;
    uint64_t p = begin & ~3ull;
    uint32_t got = 0;
    while (got < need && p < g.N) {
      const uint64_t slot = p + gl * 4;
      uint4 q = make_uint4(kEmpty, kEmpty, kEmpty, kEmpty);
      if (slot < g.N) q = *reinterpret_cast<const uint4 *>(g.pdst + slot);
      const uint32_t d[4] = {q.x, q.y, q.z, q.w};
      bool val[4];
#pragma unroll
      for (int j = 0; j < 4; j++) val[j] = d[j] != kEmpty && slot + j >= begin;
      const uint64_t B0 = __ballot(val[0]) & gmask, B1 = __ballot(val[1]) & gmask;
      const uint64_t B2 = __ballot(val[2]) & gmask, B3 = __ballot(val[3]) & gmask;
      uint32_t r = __popcll(B0 & lt) + __popcll(B1 & lt) + __popcll(B2 & lt) + __popcll(B3 & lt);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (val[j]) {
          if (got + r < need) nbrs[o + got + r] = d[j];
          r++;
        }
      }
      got += __popcll(B0) + __popcll(B1) + __popcll(B2) + __popcll(B3);
      p += 64;
    }
  }
}

__global__ void __launch_bounds__(256) k_export_tree(GView g, const uint32_t *list, uint32_t cnt,
                                                     const uint64_t *off, uint32_t *nbrs) {
  for (uint64_t w = blockIdx.x; w < cnt; w += gridDim.x) {
    uint32_t v = list[w];
    const VBlock &b = g.vb[v];
    uint32_t t = static_cast<uint32_t>(b.aux);
    uint64_t base = off[v] + kInline;
    uint64_t io = g.toff[t];
    uint32_t nb = g.tnblk[t];
    for (uint32_t j = 0; j < nb; j++) {
      uint32_t blk = g.iblk[io + j];
      uint32_t c = g.bcnt[blk];
      for (uint32_t e = threadIdx.x; e < c; e += blockDim.x) nbrs[base + e] = g.bkeys[static_cast<uint64_t>(blk) * kBlk + e];
      base += c;
    }
  }
}

// One wave: sorted neighbours of vertex v (inline, then PMA run or tree).
__global__ void k_neighbors_one(GView g, uint32_t v, uint32_t *out) {
  uint32_t lane = threadIdx.x;
  const VBlock &b = g.vb[v];
  uint32_t D = b.degree;
  uint32_t lim = D < kInline ? D : kInline;
  if (lane < lim) out[lane] = b.nbr[lane];
  if (D <= kInline) return;
  uint32_t need = D - kInline;
  uint32_t *o = out + kInline;
  if (b.aux & kAuxTree) {
    uint32_t t = static_cast<uint32_t>(b.aux);
    uint64_t io = g.toff[t];
    uint32_t nb = g.tnblk[t];
    uint64_t base = 0;
    for (uint32_t j = 0; j < nb; j++) {
      uint32_t blk = g.iblk[io + j];
      uint32_t c = g.bcnt[blk];
      for (uint32_t e = lane; e < c; e += 64) o[base + e] = g.bkeys[static_cast<uint64_t>(blk) * kBlk + e];
      base += c;
    }
    return;
  }
  uint64_t p = b.aux;
  uint32_t got = 0;
  while (got < need && p < g.N) {
    uint64_t slot = p + lane;
    uint32_t d = slot < g.N ? g.pdst[slot] : kEmpty;
    bool ok = d != kEmpty;
    uint64_t m = __ballot(ok);
    uint32_t r = __popcll(m & lanemask_lt());
    if (ok && got + r < need) o[got + r] = d;
    got += __popcll(m);
    p += 64;
  }
}

__global__ void k_init_aux(VBlock *vb, uint32_t n) {
  GRID_STRIDE(v, n) vb[v].aux = kPmaNone;
}

__global__ void k_gather_u32(const uint32_t *a, const uint32_t *flags, const uint32_t *pos, uint32_t n,
                             uint32_t *out) {
  GRID_STRIDE(i, n) if (flags[i]) out[pos[i]] = a[i];
}

__global__ void k_flag_tree_change(const Grp *grp, uint32_t ng, uint32_t *f) {
  GRID_STRIDE(x, ng) f[x] = (x == 0 || grp[x].t != grp[x - 1].t) ? 1u : 0u;
}

}  // namespace

// ===================================================================== host
Graph::Graph(uint32_t n) : n_(n) {
  if (n == 0 || n > 0x7FFFFFFFu) throw Error(-1, "vertex count must be in [1, 2^31)");
  TG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  TG_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
  for (auto &e : ev_) TG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  vb_.ensure(n, stream_);
  TG_HIP(hipMemsetAsync(vb_.get(), 0, sizeof(VBlock) * static_cast<uint64_t>(n), stream_));
  k_init_aux<<<grid_for(n, 256), 256, 0, stream_>>>(vb_.get(), n);
  TG_CHECK_LAUNCH();
  scal_.ensure(8, stream_);
  dsc_.ensure(kScalars, stream_);
  fill_u64(dsc_.get(), 0, kScalars, stream_);
  scan_scratch.lookback = true;  // single-pass scans / selects / sort on the update path
  sort_.scan.lookback = true;
  // their error words live in the batch scalars, read with the batch flags
  scan_scratch.lb.err = reinterpret_cast<uint32_t *>(dsc_.get() + kErrScan);
  sort_.scan.lb.err = reinterpret_cast<uint32_t *>(dsc_.get() + kErrSort);
  init_pma(2);
  ensure_trees(16);
  TG_HIP(hipStreamSynchronize(stream_));
}

Graph::~Graph() {
  if (side_) {
    (void)hipStreamSynchronize(side_);
    (void)hipStreamDestroy(side_);
  }
  for (auto &e : ev_)
    if (e) (void)hipEventDestroy(e);
  if (stream_) {
    (void)hipStreamSynchronize(stream_);
    (void)hipStreamDestroy(stream_);
  }
}

void Graph::fork_side(int e) {
  TG_HIP(hipEventRecord(ev_[e], stream_));
  TG_HIP(hipStreamWaitEvent(side_, ev_[e], 0));
}

void Graph::join_side(int e) {
  TG_HIP(hipEventRecord(ev_[e], side_));
  TG_HIP(hipStreamWaitEvent(stream_, ev_[e], 0));
}

GView Graph::view() const {
  if (invalid_) throw Error(-4, "graph is invalid: an earlier batch failed after it began changing the graph");
  GView g;
  g.vb = vb_.get();
  g.n = n_;
  g.pdst = pdst_.get();
  g.psrc = psrc_.get();
  g.N = pma_N_;
  g.H = pma_H_;
  g.pcnt = pcnt_.get();
  g.fence = fence_.get();
  g.bkeys = bkeys_.get();
  g.bcnt = bcnt_.get();
  g.toff = toff_.get();
  g.tnblk = tnblk_.get();
  g.iblk = iblk_.get();
  g.imin = imin_.get();
  return g;
}

void Graph::compute_bounds() {
  // density_bound (src/PMA.cc:382-394) with depth = H - l:
  //   upper = 0.75 + 0.25 * depth / H, capped at (L-1)/L - 0.001
  //   lower = 0.25 - 0.125 * depth / H
  const double limit = static_cast<double>(kLeafSlots - 1) / kLeafSlots;
  for (uint32_t l = 0; l <= pma_H_ && l < 48; l++) {
    double depth = static_cast<double>(pma_H_ - l);
    double H = pma_H_ ? static_cast<double>(pma_H_) : 1.0;
    double up = 0.75 + 0.25 * depth / H;
    if (up > limit) up = limit - 0.001;
    double lo = 0.25 - 0.125 * depth / H;
    double slots = static_cast<double>(kLeafSlots) * std::ldexp(1.0, static_cast<int>(l));
    bounds_.up[l] = static_cast<uint64_t>(std::floor(up * slots));
    // a window is underfull when count / slots <= lower (the test of
    // src/PMA.cc:1646), that is count <= floor(lower * slots)
    bounds_.lo[l] = static_cast<uint64_t>(std::floor(lo * slots));
  }
}

void Graph::init_pma(uint64_t nleaves) {
  uint32_t H = 0;
  while ((1ull << H) < nleaves) H++;
  nleaves = 1ull << H;
  pma_H_ = H;
  pma_N_ = nleaves * kLeafSlots;
  pdst_.ensure(pma_N_, stream_);
  psrc_.ensure(pma_N_, stream_);
  pcnt_.ensure(2 * nleaves, stream_);
  fence_.ensure(nleaves, stream_);
  mark_.ensure(2 * nleaves, stream_);
  fill_u32(pdst_.get(), kEmpty, pma_N_, stream_);
  fill_u32(pcnt_.get(), 0, 2 * nleaves, stream_);
  fill_u64(fence_.get(), 0, nleaves, stream_);
  TG_HIP(hipMemsetAsync(mark_.get(), 0, 2 * nleaves, stream_));
  compute_bounds();
}

void Graph::ensure_blocks(uint64_t need) {
  if (need <= bcnt_.cap) return;
  uint64_t want = need + need / 2 + 64;
  DevBuf<uint32_t> nk, nc;
  nk.ensure(want * kBlk, stream_);
  nc.ensure(want, stream_);
  if (blk_top_) {
    TG_HIP(hipMemcpyAsync(nk.get(), bkeys_.get(), blk_top_ * kBlk * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
    TG_HIP(hipMemcpyAsync(nc.get(), bcnt_.get(), blk_top_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
  }
  TG_HIP(hipStreamSynchronize(stream_));
  std::swap(bkeys_.ptr, nk.ptr);
  std::swap(bkeys_.cap, nk.cap);
  std::swap(bcnt_.ptr, nc.ptr);
  std::swap(bcnt_.cap, nc.cap);
}

void Graph::ensure_index(uint64_t need) {
  if (need <= iblk_.cap) return;
  uint64_t want = need + need / 2 + 64;
  DevBuf<uint32_t> nb, nm;
  nb.ensure(want, stream_);
  nm.ensure(want, stream_);
  if (idx_top_) {
    TG_HIP(hipMemcpyAsync(nb.get(), iblk_.get(), idx_top_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
    TG_HIP(hipMemcpyAsync(nm.get(), imin_.get(), idx_top_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
  }
  TG_HIP(hipStreamSynchronize(stream_));
  std::swap(iblk_.ptr, nb.ptr);
  std::swap(iblk_.cap, nb.cap);
  std::swap(imin_.ptr, nm.ptr);
  std::swap(imin_.cap, nm.cap);
}

void Graph::ensure_trees(uint64_t need) {
  if (need <= tnblk_.cap) return;
  uint64_t want = need + need / 2 + 16;
  DevBuf<uint64_t> no;
  DevBuf<uint32_t> nn, nv;
  no.ensure(want, stream_);
  nn.ensure(want, stream_);
  nv.ensure(want, stream_);
  if (ntrees_) {
    TG_HIP(hipMemcpyAsync(no.get(), toff_.get(), ntrees_ * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream_));
    TG_HIP(hipMemcpyAsync(nn.get(), tnblk_.get(), ntrees_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
    TG_HIP(hipMemcpyAsync(nv.get(), tvtx_.get(), ntrees_ * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
  }
  TG_HIP(hipStreamSynchronize(stream_));
  std::swap(toff_.ptr, no.ptr);
  std::swap(toff_.cap, no.cap);
  std::swap(tnblk_.ptr, nn.ptr);
  std::swap(tnblk_.cap, nn.cap);
  std::swap(tvtx_.ptr, nv.ptr);
  std::swap(tvtx_.cap, nv.cap);
}

// Compact in[flags] into out; returns the count (synchronises).
uint64_t Graph::compact_keys(const uint64_t *in, const uint32_t *flags, uint64_t n, uint64_t *out) {
  if (n == 0) return 0;
  u4_.ensure(n + 1, stream_);
  exclusive_scan_u32(flags, u4_.get(), n, u4_.get() + n, scan_scratch, stream_);
  k_scatter_flagged<<<grid_for(n, 256), 256, 0, stream_>>>(in, flags, u4_.get(), n, out);
  TG_CHECK_LAUNCH();
  return read_scalar(u4_.get() + n, stream_);
}

// Pack, validate, sort and unique the batch into k1_ without a read back:
// error bits in dsc_[0], the unique count in dsc_[1].  Returns true when
// the bucket sort ran: a kSortOverflow bit in dsc_[0] then means the
// caller must prepare again with radix = true.
bool Graph::prepare_keys_async(const uint32_t *src, const uint32_t *dst, uint64_t count, bool radix) {
  k0_.ensure(count, stream_);
  k1_.ensure(count + 1, stream_);
  if (count <= kSmallSortKeys) {
    // also clears the batch scalars
    k_pack_sort_small<<<1, 1024, 0, stream_>>>(src, dst, static_cast<uint32_t>(count), n_, k1_.get(), dsc_.get(),
                                               dsc_.get() + 1);
    TG_CHECK_LAUNCH();
    return false;  // unique keys in k1_, their count in dsc_[1]
  }
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < n_) bits++;
  fill_u64(dsc_.get(), 0, kErrSort, stream_);  // the look-back error words stay until checked
  if (count <= kBucketSortMax && !radix) {
    if (sbcnt_.ensure(kBuckets, stream_)) fill_u32(sbcnt_.get(), 0, kBuckets, stream_);
    if (sbcur_.ensure(kBuckets, stream_)) fill_u32(sbcur_.get(), 0, kBuckets, stream_);
    sboff_.ensure(kBuckets, stream_);
    sbkeys_.ensure(count, stream_);
    const uint32_t shift = 2 * bits > kBucketBits ? 2 * bits - kBucketBits : 0;
    const unsigned nb = static_cast<unsigned>((count + kBucketTile - 1) / kBucketTile);
    k_bucket_count<false><<<nb, kBucketBlock, 0, stream_>>>(src, dst, static_cast<uint32_t>(count), nullptr, nullptr,
                                                            0u, n_, bits, shift, k0_.get(), sbcnt_.get(), dsc_.get());
    TG_CHECK_LAUNCH();
    k_bucket_scatter<false><<<nb, kBucketBlock, 0, stream_>>>(k0_.get(), static_cast<uint32_t>(count), bits, shift,
                                                       sbcnt_.get(), sbcur_.get(), sboff_.get(), sbkeys_.get());
    TG_CHECK_LAUNCH();
    k_bucket_sort<<<kBuckets, kBucketBlock, 0, stream_>>>(sbkeys_.get(), static_cast<uint32_t>(count),
                                                          sboff_.get(), sbcnt_.get(), sbcur_.get(), dsc_.get());
    TG_CHECK_LAUNCH();
    select(SelUnique{sbkeys_.get(), k1_.get()}, count, dsc_.get() + 1, scan_scratch, stream_);
    return true;
  }
  {
    k_pack<<<grid_for(count, 256), 256, 0, stream_>>>(src, dst, count, n_, k0_.get(), dsc_.get());
    TG_CHECK_LAUNCH();
    radix_sort_keys(k0_.get(), count, bits, bits, sort_, stream_);
  }
  select(SelUnique{k0_.get(), k1_.get()}, count, dsc_.get() + 1, scan_scratch, stream_);
  return false;
}

// The mixed batch's sort: inserts (is, id, ni) and deletes (ds, dd, nd)
// as tagged keys, sorted, split by each edge's last op into I (k1_, count
// in dsc_[1]) and D (kD_, count in dsc_[9]); error bits in dsc_[0].
// Returns true when the bucket sort ran (see prepare_keys_async).
bool Graph::prepare_mixed_async(const uint32_t *is, const uint32_t *id, uint64_t ni, const uint32_t *ds,
                                const uint32_t *dd, uint64_t nd, bool radix) {
  const uint64_t count = ni + nd;
  k1_.ensure(count + 1, stream_);
  kD_.ensure(count + 1, stream_);
  // (batches of at most kSmallSortKeys edges are sorted by k_small_mixed_sort instead)
  k0_.ensure(count, stream_);
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < n_) bits++;
  fill_u64(dsc_.get(), 0, kErrSort, stream_);  // the look-back error words stay until checked
  const uint64_t *sorted = nullptr;
  if (count <= kBucketSortMax && !radix) {
    if (sbcnt_.ensure(kBuckets, stream_)) fill_u32(sbcnt_.get(), 0, kBuckets, stream_);
    if (sbcur_.ensure(kBuckets, stream_)) fill_u32(sbcur_.get(), 0, kBuckets, stream_);
    sboff_.ensure(kBuckets, stream_);
    sbkeys_.ensure(count, stream_);
    // compressed key (src << bits | dst) of 2 * bits bits; the op bit below
    // it does not take part in the bucket
    const uint32_t shift = 2 * bits > kBucketBits ? 2 * bits - kBucketBits : 0;
    const unsigned nb = static_cast<unsigned>((count + kBucketTile - 1) / kBucketTile);
    k_bucket_count<true><<<nb, kBucketBlock, 0, stream_>>>(is, id, static_cast<uint32_t>(ni), ds, dd,
                                                           static_cast<uint32_t>(nd), n_, bits, shift, k0_.get(),
                                                           sbcnt_.get(), dsc_.get());
    TG_CHECK_LAUNCH();
    k_bucket_scatter<true><<<nb, kBucketBlock, 0, stream_>>>(k0_.get(), static_cast<uint32_t>(count), bits, shift,
                                                             sbcnt_.get(), sbcur_.get(), sboff_.get(), sbkeys_.get());
    TG_CHECK_LAUNCH();
    k_bucket_sort<<<kBuckets, kBucketBlock, 0, stream_>>>(sbkeys_.get(), static_cast<uint32_t>(count),
                                                          sboff_.get(), sbcnt_.get(), sbcur_.get(), dsc_.get());
    TG_CHECK_LAUNCH();
    sorted = sbkeys_.get();
  } else {
    k_pack_mixed<<<grid_for(count, 256), 256, 0, stream_>>>(is, id, ni, ds, dd, nd, n_, k0_.get(), dsc_.get());
    TG_CHECK_LAUNCH();
    radix_sort_keys(k0_.get(), count, bits + 1, bits, sort_, stream_);  // dst << 1 | op, then src
    sorted = k0_.get();
  }
  select2(SelSplitOps{sorted, count, k1_.get(), kD_.get()}, count, dsc_.get() + 1, dsc_.get() + 9, scan_scratch,
          stream_);
  return count <= kBucketSortMax && !radix;
}

void Graph::trim_scratch() noexcept {
  (void)hipStreamSynchronize(stream_);
  auto trim = [](auto &b) {
    if (b.cap > kKeepScratch) b.release();
  };
  trim(k0_), trim(k1_), trim(k2_), trim(k3_), trim(kD_), trim(kE_), trim(fD_);
  trim(f0_), trim(f1_), trim(f2_), trim(f3_), trim(g0_), trim(g2_), trim(g3_);
  trim(u0_), trim(u1_), trim(u2_), trim(u3_), trim(u4_), trim(kick_), trim(route_), trim(tmp_old_);
  trim(ops_val_), trim(ops_ins_), trim(ops_del_), trim(sbkeys_), trim(sort_.alt), trim(sort_.alt_vals);
  trim(wv0_), trim(wv1_), trim(wput_), trim(dslot_), trim(dhint_);
  for (auto &b : s32_) trim(b);
  for (auto &b : s8_) trim(b);
  for (auto &b : s64_) trim(b);
}

// Internal consistency flags raised by kernels during the batch.
void Graph::check_batch_flags() {
  uint64_t h[kScalars];
  read_scalars(dsc_.get(), h, kScalars, stream_);
  check_flags(h);
}

void Graph::check_prep(const uint64_t *h) {
  if (h[kErrScan] | h[kErrSort]) {
    fill_u64(dsc_.get() + kErrSort, 0, 2, stream_);
    fill_u64(dsc_.get(), 0, 1, stream_);
    throw Error(-4, "batch preparation: a single-pass scan gave up a look-back wait (graph unchanged)");
  }
  if (h[0] & 1) {
    fill_u64(dsc_.get(), 0, 1, stream_);
    throw Error(-1, "vertex id out of range");
  }
}

template <class F>
void Graph::mutate(F &&body) {
  try {
    body();
  } catch (const std::exception &e) {
    invalid_ = true;
    throw Error(-4, std::string("batch failed after it began changing the graph; the graph is now invalid: ") +
                        e.what());
  }
}

// The same checks on batch scalars already read back.
void Graph::check_flags(const uint64_t *h) {
  uint64_t f = h[0];
  if (f) fill_u64(dsc_.get(), 0, 1, stream_);  // leave the flags clear for the next batch
  if (h[kErrScan] | h[kErrSort]) {
    fill_u64(dsc_.get() + kErrSort, 0, 2, stream_);
    throw Error(-4, "single-pass scan: a look-back wait gave up");
  }
  if (f & 2) throw Error(-4, "PMA count tree out of sync (gather)");
  if (f & 4) throw Error(-4, "PMA delete: key not found");
}

// ------------------------------------------------------------ PMA insert
uint64_t Graph::pma_leaf_counts(uint32_t *host, uint64_t cap) {
  const uint64_t nl = 1ull << pma_H_;
  const uint64_t n = cap < nl ? cap : nl;
  if (n) {
    TG_HIP(hipMemcpyAsync(host, pcnt_.get(), n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_));
    TG_HIP(hipStreamSynchronize(stream_));
  }
  return nl;
}

void Graph::rebuild_count_tree() {
  if (!tree_dirty_) return;
  tree_dirty_ = false;
  for (uint32_t L0 = 0; L0 < pma_H_; L0 += 9) {
    uint64_t nin = 1ull << (pma_H_ - L0);
    k_tree_up<<<grid_for(nin, 512, 1ull << 31), 256, 0, stream_>>>(pcnt_.get(), pma_H_, L0);
    TG_CHECK_LAUNCH();
  }
}

void Graph::pma_rebuild(const uint64_t *keys, uint64_t k, uint64_t total) {
  // gather every old element, grow, spread merge(old, keys) over the array
  uint64_t old_total = total - k;
  tmp_old_.ensure(old_total + 1, stream_);
  GView ov = view();
  uint64_t onl = 1ull << pma_H_;
  // one window covering the whole old array
  DevBuf<uint8_t> &winbuf = s8_[0];
  winbuf.ensure(sizeof(Win) + 64, stream_);
  Win w0{0, pma_H_, 0, static_cast<uint32_t>(k), static_cast<uint32_t>(old_total)};
  TG_HIP(hipMemcpyAsync(winbuf.get(), &w0, sizeof(Win), hipMemcpyHostToDevice, stream_));
  TG_HIP(hipStreamSynchronize(stream_));
  u0_.ensure(onl + 1, stream_);
  u1_.ensure(2, stream_);
  fill_u32(u1_.get(), 0, 1, stream_);  // task_base = {0}
  const Win *dw = reinterpret_cast<const Win *>(winbuf.get());
  k_task_counts<<<grid_for(onl, 256), 256, 0, stream_>>>(ov, dw, u1_.get(), 1, onl, u0_.get(), nullptr);
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(u0_.get(), u0_.get(), onl, u0_.get() + onl, scan_scratch, stream_);
  fill_u64(scal_.get(), 0, 1, stream_);
  k_gather<<<grid_for(onl * 64, 256), 256, 0, stream_>>>(ov, dw, u1_.get(), 1, onl, u0_.get(), tmp_old_.get(),
                                                         scal_.get(), nullptr);
  TG_CHECK_LAUNCH();
  if (read_scalar(scal_.get(), stream_)) throw Error(-3, "PMA count tree out of sync (gather)");
  // new geometry: density about 0.5 after the rebuild
  uint64_t nleaves = 2;
  while (nleaves * kLeafSlots / 2 < total) nleaves <<= 1;
  DevBuf<uint32_t> keep_dst, keep_src;  // old arrays released after init
  std::swap(keep_dst.ptr, pdst_.ptr);
  std::swap(keep_dst.cap, pdst_.cap);
  std::swap(keep_src.ptr, psrc_.ptr);
  std::swap(keep_src.cap, psrc_.cap);
  pcnt_.release();
  fence_.release();
  mark_.release();
  init_pma(nleaves);
  GView nv = view();
  Win w1{0, pma_H_, 0, static_cast<uint32_t>(k), static_cast<uint32_t>(old_total)};
  TG_HIP(hipMemcpyAsync(winbuf.get(), &w1, sizeof(Win), hipMemcpyHostToDevice, stream_));
  TG_HIP(hipStreamSynchronize(stream_));
  uint64_t nl = 1ull << pma_H_;
  // taskoff for the single window starts at 0: reuse u0_[0] == 0
  k_fill<<<grid_for(nl * 64, 256), 256, 0, stream_>>>(nv, dw, u1_.get(), 1, nl, u0_.get(), tmp_old_.get(), keys,
                                                     nullptr);
  TG_CHECK_LAUNCH();
  tree_dirty_ = true;
  pma_elems_ = total;
  TG_HIP(hipStreamSynchronize(stream_));
  stats_.pma_rebuilds++;
}

void Graph::pma_insert(const uint64_t *keys, uint64_t k) {
  if (k == 0) return;
  if (k > 0xFFFFFFFFull) throw Error(-1, "batch too large");
  if (pma_elems_ + k > bounds_.up[pma_H_]) {
    rebuild_count_tree();
    pma_rebuild(keys, k, pma_elems_ + k);
    return;
  }
  pma_fast_launch(keys, k);
  uint64_t h2[2];
  read_scalars(dsc_.get() + 7, h2, 2, stream_);  // fit flag, touched leaves
  if (h2[0] == 0) pma_fast_done(k, h2[1]);
  else pma_window_insert(keys, k);
}

// Touched leaves of the sorted keys (T, Tstart in f1_, f2_; their count in
// dsc_[8]) and the in-leaf merge, which writes nothing unless every touched
// leaf fits (fit flag in dsc_[7]).  Needs pma_elems_ + k within the root
// bound.
void Graph::pma_fast_launch(const uint64_t *keys, uint64_t k, const uint64_t *kdev, bool leaves_known) {
  GView g = view();
  DevBuf<uint32_t> &T = f1_, &Tstart = f2_;
  u0_.ensure(k + 1, stream_);  // leaf of key
  T.ensure(k + 1, stream_);
  Tstart.ensure(k + 1, stream_);
  uint64_t *ntp = dsc_.get() + 8, *fitp = dsc_.get() + 7;
  k_leaf_of<<<grid_for(k, 256), 256, 0, stream_>>>(g, keys, k, u0_.get(), fitp, kdev, leaves_known);
  TG_CHECK_LAUNCH();
  select(SelTouched{u0_.get(), T.get(), Tstart.get(), kdev}, k, ntp, scan_scratch, stream_);
  // kLeafLanes lanes per touched leaf (at most min(k, leaves) of them)
  const uint64_t most = k < (1ull << pma_H_) ? k : (1ull << pma_H_);
  const uint64_t wg64 = (most + kMergeGroups - 1) / kMergeGroups;
  if (wg64 > 0x7FFFFFFFull) throw Error(-1, "PMA merge: too many keys");
  const unsigned wg = static_cast<unsigned>(wg64);
  k_pma_leaf_fit<<<grid_for(k, 256), 256, 0, stream_>>>(g, bounds_.up[0], T.get(), Tstart.get(), ntp,
                                                         static_cast<uint32_t>(k), fitp, kdev);
  TG_CHECK_LAUNCH();
  k_pma_leaf_merge<<<wg, 256, 0, stream_>>>(g, T.get(), Tstart.get(), ntp, keys, fitp, dsc_.get());
  TG_CHECK_LAUNCH();
}

void Graph::pma_fast_done(uint64_t k, uint64_t nt) {
  tree_dirty_ = true;
  pma_elems_ += k;
  stats_.pma_windows += nt;
  stats_.pma_window_leaves += nt;
  stats_.pma_leaf_merges++;
}

// The window pass over the touched leaves pma_fast_launch found (a leaf did
// not fit): density windows, gather and redistribution.
void Graph::pma_window_insert(const uint64_t *keys, uint64_t k) { pma_respread(keys, k, k, false); }

// Density windows over the leaves T (f1_; their count in dsc_[8], at most
// cap), each window's share of the sorted new keys from Tstart (f2_), then
// gather and even redistribution of old and new elements.  Inserts climb to
// the lowest window within its upper bound; after deletes (low, no keys)
// the underfull leaves climb to the lowest window above its lower bound.
void Graph::pma_respread(const uint64_t *keys, uint64_t k, uint64_t cap, bool low) {
  rebuild_count_tree();
  GView g = view();
  uint64_t nl = 1ull << pma_H_;
  const uint64_t root_old = pma_elems_;
  DevBuf<uint32_t> &T = f1_, &Tstart = f2_, &chosen = f3_;
  u3_.ensure(cap + 1, stream_);  // cover, then task -> window map
  chosen.ensure(cap + 1, stream_);
  uint64_t *ntp = dsc_.get() + 8, *nwp = dsc_.get() + 9;
  if (low)
    k_climb_low<<<grid_for(cap, 256), 256, 0, stream_>>>(g, bounds_, T.get(), ntp, mark_.get(), chosen.get());
  else
    k_climb<<<grid_for(cap, 256), 256, 0, stream_>>>(g, bounds_, T.get(), Tstart.get(), ntp, mark_.get(),
                                                     chosen.get());
  TG_CHECK_LAUNCH();
  k_cover<<<grid_for(cap, 256), 256, 0, stream_>>>(g, T.get(), ntp, mark_.get(), chosen.get(), u3_.get());
  TG_CHECK_LAUNCH();
  DevBuf<uint8_t> &winbuf = route_;
  winbuf.ensure((cap + 1) * sizeof(Win) + 64, stream_);
  Win *win = reinterpret_cast<Win *>(winbuf.get());
  kick_.ensure(cap + 2, stream_);  // window leaf counts -> task_base (zeros past nw)
  fill_u32(kick_.get(), 0, cap + 1, stream_);
  select(SelOwner{g, T.get(), Tstart.get(), u3_.get(), ntp, win, kick_.get()}, cap, nwp, scan_scratch, stream_);
  k_clear_marks<<<grid_for(cap, 256), 256, 0, stream_>>>(g, T.get(), ntp, chosen.get(), mark_.get());
  TG_CHECK_LAUNCH();
  uint32_t *task_base = kick_.get();
  // scanning cap entries (zeros past nw) keeps nw on the device: task_base[w]
  // for w >= nw equals the task total, so window searches over cap are exact
  exclusive_scan_u32(task_base, task_base, cap, task_base + cap, scan_scratch, stream_);
  k_u32_to_u64<<<1, 1, 0, stream_>>>(task_base + cap, dsc_.get() + 10);
  TG_CHECK_LAUNCH();
  uint64_t h[3];
  read_scalars(dsc_.get() + 8, h, 3, stream_);  // nt, nw, ntasks
  uint32_t nw = static_cast<uint32_t>(h[1]);
  uint64_t ntasks = h[2];
  if (ntasks > nl) throw Error(-4, "PMA window tasks exceed leaves");
  u4_.ensure(ntasks + 1, stream_);
  u3_.ensure(ntasks + 1, stream_);
  uint32_t *tw = u3_.get();  // cover[] is no longer needed
  k_task_map<<<grid_for(nw, 256), 256, 0, stream_>>>(task_base, nw, tw);
  TG_CHECK_LAUNCH();
  k_task_counts<<<grid_for(ntasks, 256), 256, 0, stream_>>>(g, win, task_base, nw, ntasks, u4_.get(), tw);
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(u4_.get(), u4_.get(), ntasks, u4_.get() + ntasks, scan_scratch, stream_);
  tmp_old_.ensure(root_old + 1, stream_);  // the windows hold at most every element
  k_gather<<<grid_for(ntasks * 64, 256), 256, 0, stream_>>>(g, win, task_base, nw, ntasks, u4_.get(),
                                                            tmp_old_.get(), dsc_.get(), tw);
  TG_CHECK_LAUNCH();
  k_fill<<<grid_for(ntasks * 64, 256), 256, 0, stream_>>>(g, win, task_base, nw, ntasks, u4_.get(), tmp_old_.get(),
                                                          keys, tw);
  TG_CHECK_LAUNCH();
  tree_dirty_ = true;
  pma_elems_ += k;
  stats_.pma_windows += nw;
  stats_.pma_window_leaves += ntasks;
  if (low) stats_.pma_low_windows += nw;
}

// After deletes left a leaf at or below the leaf lower bound: an array whose
// root is at or below its own lower bound shrinks (half_list,
// src/PMA.cc:722-813; here straight to the size a rebuild gives the count,
// density between 1/4 and 1/2, never below two leaves); otherwise every
// underfull leaf's window is respread.  The gather checks each leaf's
// count against its slots, as on inserts.
void Graph::pma_rebalance_low() {
  const uint64_t nl = 1ull << pma_H_;
  if (pma_elems_ == 0) {
    if (nl > 2) {
      pdst_.release();
      psrc_.release();
      pcnt_.release();
      fence_.release();
      mark_.release();
      init_pma(2);
      tree_dirty_ = false;  // a fresh array's counts are all zero
      stats_.pma_shrinks++;
    }
    return;
  }
  if (nl > 2 && pma_elems_ <= bounds_.lo[pma_H_]) {
    rebuild_count_tree();
    pma_rebuild(nullptr, 0, pma_elems_);
    stats_.pma_shrinks++;
    return;
  }
  rebuild_count_tree();
  DevBuf<uint32_t> &T = f1_, &Tstart = f2_;
  T.ensure(nl + 1, stream_);
  Tstart.ensure(nl + 1, stream_);
  fill_u32(Tstart.get(), 0, nl + 1, stream_);
  select(SelLowLeaf{pcnt_.get(), static_cast<uint32_t>(bounds_.lo[0]), T.get(), Tstart.get()}, nl,
         dsc_.get() + 8, scan_scratch, stream_);
  const uint64_t nt = read_scalar(dsc_.get() + 8, stream_);
  if (nt == 0) return;
  // no new keys: k_fill reads nothing from the key array when a window's key
  // range is empty, any valid pointer serves
  tmp_old_.ensure(pma_elems_ + 1, stream_);
  pma_respread(tmp_old_.get(), 0, nt, true);
}

// Delete up to cap present keys (kdev: their count on the device); the
// caller lowers pma_elems_ by the count.
void Graph::pma_delete_launch(const uint64_t *keys, uint64_t cap, const uint64_t *kdev, const uint64_t *hint) {
  GView g = view();
  k_pma_delete<<<grid_for(cap, 256), 256, 0, stream_>>>(g, keys, cap, dsc_.get(), kdev,
                                                         static_cast<uint32_t>(bounds_.lo[0]),
                                                         dsc_.get() + kLowFlag, hint);
  TG_CHECK_LAUNCH();
  tree_dirty_ = true;
}

// ------------------------------------------------------------ tree update
// Launch the in-place tree update (status in dsc_[13]: 0 = done).
// Up to min(cap, kTreeSmallKeys) keys (kdev: their count on the device;
// more keys give status 1); the groups kernel lists the blocks to rewrite,
// the apply kernel's grid covers the most groups cap keys can make.
void Graph::tree_small_launch(const uint64_t *keys, uint64_t cap, bool insert, const uint64_t *kdev,
                              const uint64_t *gate, uint64_t *status, hipStream_t s) {
  GView g = view();
  DevBuf<uint32_t> &gl = s32_[30];
  const uint64_t most = cap < kTreeSmallKeys ? cap : kTreeSmallKeys;
  gl.ensure(1 + 3 * most, stream_);  // presized on the update path: no reallocation under way
  const uint32_t kk = static_cast<uint32_t>(most);
  const unsigned grid = static_cast<unsigned>(most == 0 ? 1 : (most < 1024 ? most : 1024));
  uint64_t *st = status ? status : dsc_.get() + 13;
  if (!s) s = stream_;
  if (insert) {
    k_tree_small_groups<true><<<1, 1024, 0, s>>>(g, keys, kk, gl.get(), st, kdev, gate);
    TG_CHECK_LAUNCH();
    k_tree_small_apply<true><<<grid, 256, 0, s>>>(g, keys, gl.get());
  } else {
    k_tree_small_groups<false><<<1, 1024, 0, s>>>(g, keys, kk, gl.get(), st, kdev, gate);
    TG_CHECK_LAUNCH();
    k_tree_small_apply<false><<<grid, 256, 0, s>>>(g, keys, gl.get());
  }
  TG_CHECK_LAUNCH();
}

bool Graph::tree_small(const uint64_t *keys, uint64_t k, bool insert) {
  if (k == 0 || k > kTreeSmallKeys) return false;
  tree_small_launch(keys, k, insert);
  if (read_scalar(dsc_.get() + 13, stream_) != 0) return false;
  stats_.tree_small_updates++;
  return true;
}

void Graph::tree_insert(const uint64_t *keys, uint64_t k) {
  if (k == 0) return;
  if (tree_small(keys, k, true)) return;
  tree_insert_cow(keys, k);
}

// The copy-on-write tree insert: groups, new blocks, rebuilt index segments.
void Graph::tree_insert_cow(const uint64_t *keys, uint64_t k) {
  GView g = view();
  g3_.ensure(k + 1, stream_);
  u0_.ensure(k + 1, stream_);
  u1_.ensure(k + 1, stream_);
  k_tree_loc<<<grid_for(k, 256), 256, 0, stream_>>>(g, keys, k, g3_.get());
  TG_CHECK_LAUNCH();
  k_flag_change_u64<<<grid_for(k, 256), 256, 0, stream_>>>(g3_.get(), k, u0_.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(u0_.get(), u1_.get(), k, u1_.get() + k, scan_scratch, stream_);
  uint32_t ng = read_scalar(u1_.get() + k, stream_);
  DevBuf<uint8_t> &grpbuf = s8_[1];
  grpbuf.ensure(static_cast<uint64_t>(ng + 1) * sizeof(Grp), stream_);
  Grp *grp = reinterpret_cast<Grp *>(grpbuf.get());
  DevBuf<uint32_t> &gkb = s32_[20], &ob = s32_[21], &ho = s32_[22], &gflag = s32_[23], &gpos = s32_[24],
                   &grun = s32_[25];
  gkb.ensure(ng + 1, stream_);
  ob.ensure(ng + 1, stream_);
  ho.ensure(ng + 1, stream_);
  k_make_groups<<<grid_for(k, 256), 256, 0, stream_>>>(g, g3_.get(), u0_.get(), u1_.get(), k, ng, grp, gkb.get());
  TG_CHECK_LAUNCH();
  k_group_sizes<<<grid_for(ng, 256), 256, 0, stream_>>>(g, grp, gkb.get(), ng, true, ob.get(), ho.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(ob.get(), ob.get(), ng, ob.get() + ng, scan_scratch, stream_);
  exclusive_scan_u32(ho.get(), ho.get(), ng, ho.get() + ng, scan_scratch, stream_);
  gflag.ensure(ng + 1, stream_);
  gpos.ensure(ng + 1, stream_);
  grun.ensure(ng + 1, stream_);
  k_flag_tree_change<<<grid_for(ng, 256), 256, 0, stream_>>>(grp, ng, gflag.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(gflag.get(), gpos.get(), ng, gpos.get() + ng, scan_scratch, stream_);
  // out-blocks and touched-tree runs: one read
  k_u32_pair<<<1, 1, 0, stream_>>>(ob.get() + ng, gpos.get() + ng, dsc_.get() + 11);
  TG_CHECK_LAUNCH();
  uint64_t h2[2];
  read_scalars(dsc_.get() + 11, h2, 2, stream_);
  uint64_t nout = h2[0];
  uint32_t ntr = static_cast<uint32_t>(h2[1]);
  uint64_t blk_base = blk_top_;
  ensure_blocks(blk_top_ + nout);
  blk_top_ += nout;
  g = view();
  k_tree_fill_insert<<<grid_for(nout * 256, 256, 1u << 16), 256, 0, stream_>>>(g, grp, ob.get(), ng, nout,
                                                                                blk_base, keys);
  TG_CHECK_LAUNCH();
  stats_.tree_blocks_written += nout;
  // index rebuild for the touched trees
  k_grp_run<<<grid_for(ng, 256), 256, 0, stream_>>>(gflag.get(), gpos.get(), ng, grun.get());
  TG_CHECK_LAUNCH();
  DevBuf<uint8_t> &trbuf = s8_[2];
  trbuf.ensure(static_cast<uint64_t>(ntr + 1) * sizeof(TreeRun), stream_);
  TreeRun *tr = reinterpret_cast<TreeRun *>(trbuf.get());
  DevBuf<uint32_t> &trs = s32_[26], &nbn = s32_[27], &nbo = s32_[28];
  trs.ensure(ntr + 1, stream_);
  nbn.ensure(ntr + 1, stream_);
  nbo.ensure(ntr + 1, stream_);
  k_tree_runs<<<grid_for(ng, 256), 256, 0, stream_>>>(g, grp, gflag.get(), gpos.get(), ng, ntr, tr, trs.get());
  TG_CHECK_LAUNCH();
  k_tree_run_sizes<<<grid_for(ntr, 256), 256, 0, stream_>>>(g, grp, trs.get(), ntr, ob.get(), ho.get(), tr,
                                                            nbn.get(), nbo.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(nbn.get(), nbn.get(), ntr, nbn.get() + ntr, scan_scratch, stream_);
  exclusive_scan_u32(nbo.get(), nbo.get(), ntr, nbo.get() + ntr, scan_scratch, stream_);
  k_u32_pair<<<1, 1, 0, stream_>>>(nbn.get() + ntr, nbo.get() + ntr, dsc_.get() + 11);
  TG_CHECK_LAUNCH();
  read_scalars(dsc_.get() + 11, h2, 2, stream_);
  uint64_t nidx_new = h2[0];
  uint64_t nidx_old = h2[1];
  uint64_t idx_base = idx_top_;
  ensure_index(idx_top_ + nidx_new);
  idx_top_ += nidx_new;
  g = view();
  if (nidx_old)
    k_tree_idx_copy<<<grid_for(nidx_old, 256), 256, 0, stream_>>>(g, tr, nbo.get(), ntr, nidx_old, grp, ob.get(),
                                                                  ho.get(), nbn.get(), idx_base);
  TG_CHECK_LAUNCH();
  k_tree_idx_groups<<<grid_for(ng, 256), 256, 0, stream_>>>(g, tr, ntr, grp, grun.get(), ng, ob.get(), ho.get(),
                                                            nbn.get(), idx_base, blk_base);
  TG_CHECK_LAUNCH();
  k_tree_commit<<<grid_for(ntr, 256), 256, 0, stream_>>>(g, tr, ntr, nbn.get(), idx_base);
  TG_CHECK_LAUNCH();
}

void Graph::tree_delete(const uint64_t *keys, uint64_t k) {
  if (k == 0) return;
  if (tree_small(keys, k, false)) return;
  tree_delete_cow(keys, k);
}

// The copy-on-write tree delete.
void Graph::tree_delete_cow(const uint64_t *keys, uint64_t k) {
  GView g = view();
  g3_.ensure(k + 1, stream_);
  u0_.ensure(k + 1, stream_);
  u1_.ensure(k + 1, stream_);
  k_tree_loc<<<grid_for(k, 256), 256, 0, stream_>>>(g, keys, k, g3_.get());
  TG_CHECK_LAUNCH();
  k_flag_change_u64<<<grid_for(k, 256), 256, 0, stream_>>>(g3_.get(), k, u0_.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(u0_.get(), u1_.get(), k, u1_.get() + k, scan_scratch, stream_);
  uint32_t ng = read_scalar(u1_.get() + k, stream_);
  DevBuf<uint8_t> &grpbuf = s8_[1];
  grpbuf.ensure(static_cast<uint64_t>(ng + 1) * sizeof(Grp), stream_);
  Grp *grp = reinterpret_cast<Grp *>(grpbuf.get());
  DevBuf<uint32_t> &gkb = s32_[20], &ob = s32_[21], &ho = s32_[22], &gflag = s32_[23], &gpos = s32_[24],
                   &grun = s32_[25];
  gkb.ensure(ng + 1, stream_);
  ob.ensure(ng + 1, stream_);
  ho.ensure(ng + 1, stream_);
  k_make_groups<<<grid_for(k, 256), 256, 0, stream_>>>(g, g3_.get(), u0_.get(), u1_.get(), k, ng, grp, gkb.get());
  TG_CHECK_LAUNCH();
  k_group_sizes<<<grid_for(ng, 256), 256, 0, stream_>>>(g, grp, gkb.get(), ng, false, ob.get(), ho.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(ob.get(), ob.get(), ng, ob.get() + ng, scan_scratch, stream_);
  exclusive_scan_u32(ho.get(), ho.get(), ng, ho.get() + ng, scan_scratch, stream_);
  gflag.ensure(ng + 1, stream_);
  gpos.ensure(ng + 1, stream_);
  grun.ensure(ng + 1, stream_);
  k_flag_tree_change<<<grid_for(ng, 256), 256, 0, stream_>>>(grp, ng, gflag.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(gflag.get(), gpos.get(), ng, gpos.get() + ng, scan_scratch, stream_);
  // out-blocks and touched-tree runs: one read
  k_u32_pair<<<1, 1, 0, stream_>>>(ob.get() + ng, gpos.get() + ng, dsc_.get() + 11);
  TG_CHECK_LAUNCH();
  uint64_t h2[2];
  read_scalars(dsc_.get() + 11, h2, 2, stream_);
  uint64_t nout = h2[0];
  uint32_t ntr = static_cast<uint32_t>(h2[1]);
  uint64_t blk_base = blk_top_;
  ensure_blocks(blk_top_ + nout);
  blk_top_ += nout;
  g = view();
  k_tree_fill_delete<<<grid_for(static_cast<uint64_t>(ng) * 256, 256, 1u << 16), 256, 0, stream_>>>(
      g, grp, ob.get(), ng, blk_base, keys);
  TG_CHECK_LAUNCH();
  stats_.tree_blocks_written += nout;
  k_grp_run<<<grid_for(ng, 256), 256, 0, stream_>>>(gflag.get(), gpos.get(), ng, grun.get());
  TG_CHECK_LAUNCH();
  DevBuf<uint8_t> &trbuf = s8_[2];
  trbuf.ensure(static_cast<uint64_t>(ntr + 1) * sizeof(TreeRun), stream_);
  TreeRun *tr = reinterpret_cast<TreeRun *>(trbuf.get());
  DevBuf<uint32_t> &trs = s32_[26], &nbn = s32_[27], &nbo = s32_[28];
  trs.ensure(ntr + 1, stream_);
  nbn.ensure(ntr + 1, stream_);
  nbo.ensure(ntr + 1, stream_);
  k_tree_runs<<<grid_for(ng, 256), 256, 0, stream_>>>(g, grp, gflag.get(), gpos.get(), ng, ntr, tr, trs.get());
  TG_CHECK_LAUNCH();
  k_tree_run_sizes<<<grid_for(ntr, 256), 256, 0, stream_>>>(g, grp, trs.get(), ntr, ob.get(), ho.get(), tr,
                                                            nbn.get(), nbo.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(nbn.get(), nbn.get(), ntr, nbn.get() + ntr, scan_scratch, stream_);
  exclusive_scan_u32(nbo.get(), nbo.get(), ntr, nbo.get() + ntr, scan_scratch, stream_);
  k_u32_pair<<<1, 1, 0, stream_>>>(nbn.get() + ntr, nbo.get() + ntr, dsc_.get() + 11);
  TG_CHECK_LAUNCH();
  read_scalars(dsc_.get() + 11, h2, 2, stream_);
  uint64_t nidx_new = h2[0];
  uint64_t nidx_old = h2[1];
  uint64_t idx_base = idx_top_;
  ensure_index(idx_top_ + nidx_new + 1);
  idx_top_ += nidx_new;
  g = view();
  if (nidx_old)
    k_tree_idx_copy<<<grid_for(nidx_old, 256), 256, 0, stream_>>>(g, tr, nbo.get(), ntr, nidx_old, grp, ob.get(),
                                                                  ho.get(), nbn.get(), idx_base);
  TG_CHECK_LAUNCH();
  k_tree_idx_groups<<<grid_for(ng, 256), 256, 0, stream_>>>(g, tr, ntr, grp, grun.get(), ng, ob.get(), ho.get(),
                                                            nbn.get(), idx_base, blk_base);
  TG_CHECK_LAUNCH();
  k_tree_commit<<<grid_for(ntr, 256), 256, 0, stream_>>>(g, tr, ntr, nbn.get(), idx_base);
  TG_CHECK_LAUNCH();
}

// ------------------------------------------------------------ batches
void Graph::insert_batch(const uint32_t *src, const uint32_t *dst, uint64_t count) {
  if (count == 0) return;
  TrimAfter trim_after{*this, count > kKeepScratch};
  if (ws_) throw Error(-1, "weighted graph: insert with weights");
  if (count > 0xFFFFFFFFull) throw Error(-1, "batch larger than 2^32 edges");
  if (read_only_) throw Error(-1, "graph is a query-only snapshot");
  // sort, unique, existence and source runs sized by count; the unique count
  // stays on the device until the one read below
  const uint64_t *R = nullptr;
  const bool small = count <= kSmallSortKeys;
  const bool fused = count <= kBucketSortMax;  // the insert launched before the read
  bool pma_launched = false;
  uint64_t h[kScalars];
  // a bucket sort that overflowed is redone by radix (second pass)
  for (int pass = 0; pass < 2; pass++) {
    const bool bucket = prepare_keys_async(src, dst, count, pass == 1);
    R = k1_.get();  // after the prepare, which sizes k1_
    GView g = view();
    f0_.ensure(count + 1, stream_);
    f1_.ensure(count + 1, stream_);
    s32_[0].ensure(count + 1, stream_);
    if (small || count > kBucketSortMax) {
      k_exists2_n<<<grid_for(count, 256), 256, 0, stream_>>>(g, R, dsc_.get() + 1, count, f0_.get(), f1_.get());
      TG_CHECK_LAUNCH();
    }
    if (small) {
      // new-key prefix, source runs and run_beg[nruns] = nR in one workgroup
      k_small_runs<true><<<1, 1024, 0, stream_>>>(R, f1_.get(), static_cast<uint32_t>(count), nullptr,
                                                  s32_[0].get(), dsc_.get());
      TG_CHECK_LAUNCH();
    } else {
      // new-key prefix and source runs in one launch; up to kBucketSortMax
      // keys it also does the lookups, one key per thread
      if (count <= kBucketSortMax) {
        const uint32_t nb = static_cast<uint32_t>((count + kPrepBlock - 1) / kPrepBlock);
        const uint32_t epoch = scan_scratch.lb.begin(2ull * nb, stream_);
        k_insert_prep<1, true><<<nb, kPrepBlock, 0, stream_>>>(
            g, R, dsc_.get() + 1, count, f0_.get(), f1_.get(), s32_[0].get(), dsc_.get() + 2,
            scan_scratch.lb.state.get(), scan_scratch.lb.misc.get(), scan_scratch.lb.err_word(), nb, epoch);
      } else {
        const uint32_t nb = static_cast<uint32_t>((count + 8 * kPrepBlock - 1) / (8 * kPrepBlock));
        const uint32_t epoch = scan_scratch.lb.begin(2ull * nb, stream_);
        k_insert_prep<8, false><<<nb, kPrepBlock, 0, stream_>>>(
            g, R, dsc_.get() + 1, count, f0_.get(), f1_.get(), s32_[0].get(), dsc_.get() + 2,
            scan_scratch.lb.state.get(), scan_scratch.lb.misc.get(), scan_scratch.lb.err_word(), nb, epoch);
      }
      TG_CHECK_LAUNCH();
    }
    if (debug_fail & 1u) {
      debug_fail &= ~1u;
      fill_u64(dsc_.get() + kErrScan, 1, 1, stream_);
    }
    if (fused) {
      // the insert on the device's counts (dsc_[1], [2]), gated on the
      // preparation's error words, before the one read
      presize_insert(count, static_cast<uint32_t>(count));
      mutate([&] { pma_launched = insert_launch(R, count, static_cast<uint32_t>(count), dsc_.get() + 1, dsc_.get()); });
    }
    // counts, error bits and the look-back error words in one read
    read_scalars(dsc_.get(), h, kScalars, stream_);
    if (!bucket || !(h[0] & kSortOverflow)) break;
  }
  if (fused && h[bsw::kGate + 2] != 1) {
    check_prep(h);  // the gate closed on a preparation error: nothing was written
    throw Error(-4, "insert batch: the update gate closed without an error word");
  }
  if (!fused) check_prep(h);
  stats_.last_ins_keys = h[1];
  stats_.last_ins_sources = h[2];
  stats_.sum_ins_sources += h[2];
  stats_.last_del_keys = stats_.last_del_sources = 0;
  if (fused) {
    mutate([&] {
      if (insert_finish(h, R, pma_launched)) check_flags(h);
    });
    return;
  }
  if (h[1] == 0) return;
  // run_beg[nruns] = nR is written; the counts stay in dsc_[1], [2]
  insert_runs(R, h[1], static_cast<uint32_t>(h[2]), dsc_.get() + 1);
}

// Insert sorted unique keys (R must not alias the batch scratch buffers).
void Graph::insert_sorted(const uint64_t *R, uint64_t nR) {
  if (nR == 0) return;
  if (read_only_) throw Error(-1, "graph is a query-only snapshot");
  GView g = view();
  // existence + new-key prefix
  f0_.ensure(nR + 1, stream_);
  f1_.ensure(nR + 1, stream_);
  uint32_t *ex = f0_.get(), *nidx = f1_.get();
  k_exists2<<<grid_for(nR, 256), 256, 0, stream_>>>(g, R, nR, ex, nidx);
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(nidx, nidx, nR, nidx + nR, scan_scratch, stream_);
  // source runs (count in dsc_[2]; the key count goes to dsc_[1] beside it)
  DevBuf<uint32_t> &run_beg = s32_[0];
  run_beg.ensure(nR + 1, stream_);
  select(SelRunStart{R, run_beg.get()}, nR, dsc_.get() + 2, scan_scratch, stream_);
  k_put_u32<<<1, 1, 0, stream_>>>(run_beg.get(), dsc_.get() + 2, static_cast<uint32_t>(nR));
  TG_CHECK_LAUNCH();
  fill_u64(dsc_.get() + 1, nR, 1, stream_);
  uint64_t h[kScalars];
  read_scalars(dsc_.get(), h, kScalars, stream_);
  check_prep(h);
  insert_runs(R, nR, static_cast<uint32_t>(h[2]), dsc_.get() + 1);
}

// Scratch of the insert's common path, bounded by the batch: sized before
// the first write.  The fallbacks (promotion, the copy-on-write tree pass,
// the PMA window pass and growth) size theirs by counts known only later.
void Graph::presize_insert(uint64_t nR, uint32_t nruns) {
  for (int i : {5, 6, 7}) s32_[i].ensure(nruns + 1, stream_);
  k3_.ensure(nR + 1, stream_);
  g0_.ensure(nR + 1, stream_);
  u0_.ensure(nR + 1, stream_);
  f2_.ensure(nR + 1, stream_);  // f1_ (the new-key prefix, live here) already holds nR + 1
  s32_[30].ensure(1 + 3 * std::min<uint64_t>(nR, kTreeSmallKeys), stream_);
}

void Graph::insert_runs(const uint64_t *R, uint64_t nR, uint32_t nruns, const uint64_t *cw) {
  presize_insert(nR, nruns);
  mutate([&] {
    const bool pma_launched = insert_launch(R, nR, nruns, cw, nullptr);
    uint64_t hs[kScalars];
    read_scalars(dsc_.get(), hs, kScalars, stream_);
    if (insert_finish(hs, R, pma_launched)) check_flags(hs);
  });
}

// The insert proper: existence flags in f0_, the new-key prefix in f1_ and
// the run starts (with run_beg[nruns] = nR) in s32_[0] are ready; the key
// and run counts are on the device at cw.  Every kernel is sized by the
// capacities and reads the counts k_inline_merge published, so nothing
// here waits for the host.
bool Graph::insert_launch(const uint64_t *R, uint64_t capR, uint32_t capRuns, const uint64_t *cw,
                          const uint64_t *gate, const uint32_t *ileaf) {
  lists_dirty_ = true;
  ++epoch_;
  GView g = view();
  DevBuf<uint32_t> &run_beg = s32_[0], &promote = s32_[5], &ppos = s32_[6], &old_deg = s32_[7];
  promote.ensure(capRuns + 1, stream_);
  ppos.ensure(capRuns + 1, stream_);
  old_deg.ensure(capRuns + 1, stream_);  // promotions need the degrees before the merge
  k3_.ensure(capR + 1, stream_);
  u0_.ensure(capR + 1, stream_);
  g0_.ensure(capR + 1, stream_);
  // the inline merge of every run and the kicked keys split by tier (PMA
  // keys into k3_, count dsc_[3], their leaves into u0_; tree keys into g0_,
  // [4]; promotions [5])
  const uint32_t nb = static_cast<uint32_t>(std::max<uint64_t>(1, (capRuns + kRunBlock - 1) / kRunBlock));
  const uint32_t epoch = scan_scratch.lb.begin(3ull * nb, stream_);
  k_insert_runs<<<nb, kRunBlock, 0, stream_>>>(g, R, f0_.get(), f1_.get(), ileaf, run_beg.get(), cw, gate,
                                               dsc_.get() + bsw::kGate, capR, capRuns, k3_.get(), u0_.get(), g0_.get(),
                                               promote.get(), ppos.get(), old_deg.get(), dsc_.get() + 3,
                                               scan_scratch.lb.state.get(), scan_scratch.lb.misc.get(),
                                               scan_scratch.lb.err_word(), nb, epoch);
  TG_CHECK_LAUNCH();
  if (debug_fail & 2u) {
    debug_fail &= ~2u;
    throw Error(-2, "injected allocation failure after the inline merge");
  }
  // The in-place tree update and the in-leaf PMA merge (disjoint data) run
  // on the device's counts.  The tree update waits (status 2) while
  // promotions are pending: their trees must exist first.  The PMA merge is
  // launched when every key could go to the PMA within the root bound (else
  // the general insert decides after the read).
  // A large batch runs the tree update on the side stream, beside the PMA
  // kernels, and joins it back before anything after the half.
  const bool pma_launched = pma_elems_ + capR <= bounds_.up[pma_H_];
  const bool side = capR >= kSideMin;
  if (side) fork_side(0);
  tree_small_launch(g0_.get(), capR, true, dsc_.get() + 4, dsc_.get() + 5, nullptr, side ? side_ : nullptr);
  if (pma_launched) pma_fast_launch(k3_.get(), capR, dsc_.get() + 3, true);
  if (side) join_side(1);
  return pma_launched;
}

bool Graph::insert_finish(const uint64_t *hs, const uint64_t *R, bool pma_launched) {
  const uint32_t nruns = static_cast<uint32_t>(hs[bsw::kGate + 1]);
  const uint64_t npma = hs[3];
  uint64_t ntree = hs[4];
  const uint32_t np = static_cast<uint32_t>(hs[5]);
  bool tree_ok = ntree == 0 || hs[13] == 0;
  const bool pma_ok = npma == 0 || (pma_launched && hs[7] == 0);
  if (ntree && tree_ok) stats_.tree_small_updates++;
  if (npma && pma_ok) pma_fast_done(npma, hs[8]);
  if (np == 0 && tree_ok && pma_ok) return true;
  GView g = view();
  DevBuf<uint32_t> &run_beg = s32_[0], &promote = s32_[5], &ppos = s32_[6], &old_deg = s32_[7];
  // promotions: new trees for vertices routed to the B-tier the first time
  // (their PMA runs are read after the in-leaf merge, which keeps each run
  // contiguous and its begin slot exact)
  uint64_t next = 0;
  DevBuf<uint64_t> &ext = s64_[0];
  if (np) {
    DevBuf<uint32_t> &pv = s32_[8], &pcount = s32_[9], &pext = s32_[10], &pod = s32_[11];
    pv.ensure(np + 1, stream_);
    pcount.ensure(np + 1, stream_);
    pext.ensure(np + 1, stream_);
    pod.ensure(np + 1, stream_);
    k_promote_info<<<grid_for(nruns, 256), 256, 0, stream_>>>(R, run_beg.get(), promote.get(), ppos.get(), nruns, g,
                                                              pv.get(), pcount.get());
    TG_CHECK_LAUNCH();
    k_gather_u32<<<grid_for(nruns, 256), 256, 0, stream_>>>(old_deg.get(), promote.get(), ppos.get(), nruns,
                                                           pod.get());
    TG_CHECK_LAUNCH();
    k_promote_old_counts<<<grid_for(np, 256), 256, 0, stream_>>>(pv.get(), np, pod.get(), pcount.get());
    TG_CHECK_LAUNCH();
    exclusive_scan_u32(pcount.get(), pext.get(), np, pext.get() + np, scan_scratch, stream_);
    next = read_scalar(pext.get() + np, stream_);
    ext.ensure(next + 1, stream_);
    if (next) {
      // the runs leave the PMA as they are copied out
      fill_u64(dsc_.get() + kLowFlag, 0, 1, stream_);
      k_extract_pma<<<grid_for(static_cast<uint64_t>(np) * 64, 256), 256, 0, stream_>>>(
          g, pv.get(), pcount.get(), pext.get(), np, ext.get(), static_cast<uint32_t>(bounds_.lo[0]),
          dsc_.get() + kLowFlag);
      TG_CHECK_LAUNCH();
      pma_elems_ -= next;
      tree_dirty_ = true;
    }
    ensure_trees(ntrees_ + np);
    g = view();
    k_set_tree<<<grid_for(np, 256), 256, 0, stream_>>>(g, pv.get(), np, ntrees_, tvtx_.get());
    TG_CHECK_LAUNCH();
    ntrees_ += np;
    stats_.promotions += np;
  }
  if (next) {
    g0_.ensure(ntree + next + 1, stream_, true, ntree);
    TG_HIP(hipMemcpyAsync(g0_.get() + ntree, ext.get(), next * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream_));
    ntree += next;
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < n_) bits++;
    radix_sort_keys(g0_.get(), ntree, bits, bits, sort_, stream_);
  }
  if (!tree_ok || next) {
    // waited for promotions (status 2) or did not fit in place (status 1):
    // copy-on-write.  After a promotion the in-place update cannot succeed:
    // each new tree has keys of this batch and no blocks yet.
    tree_insert_cow(g0_.get(), ntree);
  }
  if (!pma_ok) {
    if (pma_launched) pma_window_insert(k3_.get(), npma);
    else pma_insert(k3_.get(), npma);  // may be beyond the root bound: the general insert decides
  }
  check_batch_flags();
  // the promoted runs left the PMA: respread what they left underfull
  if (next && read_scalar(dsc_.get() + kLowFlag, stream_)) pma_rebalance_low();
  return false;
}

void Graph::delete_batch(const uint32_t *src, const uint32_t *dst, uint64_t count) {
  if (count == 0) return;
  TrimAfter trim_after{*this, count > kKeepScratch};
  if (count > 0xFFFFFFFFull) throw Error(-1, "batch larger than 2^32 edges");
  if (read_only_) throw Error(-1, "graph is a query-only snapshot");
  // sort, unique, the existing keys and their source runs sized by count;
  // the three counts come back in one read
  const bool small = count <= kSmallSortKeys;
  const bool fused = count <= kBucketSortMax;  // the delete launched before the read
  uint64_t *dsd = dsc_.get() + bsw::kDel;
  uint64_t h[kScalars];
  // a bucket sort that overflowed is redone by radix (second pass)
  for (int pass = 0; pass < 2; pass++) {
    const bool bucket = prepare_keys_async(src, dst, count, pass == 1);
    GView g = view();
    k2_.ensure(count + 1, stream_);
    f0_.ensure(count + 1, stream_);
    if (small || count > kBucketSortMax) {
      k_exists_n<<<grid_for(count, 256), 256, 0, stream_>>>(g, k1_.get(), dsc_.get() + 1, count, f0_.get());
      TG_CHECK_LAUNCH();
    }
    s32_[0].ensure(count + 1, stream_);
    if (small) {
      // existing keys, their source runs and run_beg[nruns] in one workgroup
      k_small_runs<false><<<1, 1024, 0, stream_>>>(k1_.get(), f0_.get(), static_cast<uint32_t>(count), k2_.get(),
                                                   s32_[0].get(), dsc_.get());
      TG_CHECK_LAUNCH();
    } else {
      if (count <= kBucketSortMax) {
        // existence and compaction of the existing keys in one launch
        const uint32_t nb = static_cast<uint32_t>((count + kPrepBlock - 1) / kPrepBlock);
        const uint32_t epoch = scan_scratch.lb.begin(nb, stream_);
        k_delete_prep<<<nb, kPrepBlock, 0, stream_>>>(g, k1_.get(), dsc_.get() + 1, k2_.get(), dsc_.get() + 2,
                                                      scan_scratch.lb.state.get(), scan_scratch.lb.misc.get(),
                                                      scan_scratch.lb.err_word(), nb, epoch);
        TG_CHECK_LAUNCH();
      } else {
        select(SelFlagged{k1_.get(), f0_.get(), k2_.get()}, count, dsc_.get() + 2, scan_scratch, stream_);
      }
      select(SelRunStartN{k2_.get(), s32_[0].get(), dsc_.get() + 2}, count, dsc_.get() + 3, scan_scratch,
             stream_);
    }
    if (debug_fail & 1u) {
      debug_fail &= ~1u;
      fill_u64(dsc_.get() + kErrScan, 1, 1, stream_);
    }
    if (fused) {
      // the delete on the device's counts (dsc_[2], [3]), gated on the
      // preparation's error words, before the one read
      presize_delete(count, static_cast<uint32_t>(count));
      mutate([&] {
        delete_launch(k2_.get(), nullptr, nullptr, count, static_cast<uint32_t>(count), s32_[0].get(),
                      dsc_.get() + 2, dsc_.get(), false, false, dsd);
      });
    }
    read_scalars(dsc_.get(), h, kScalars, stream_);
    if (!bucket || !(h[0] & kSortOverflow)) break;
  }
  if (fused && h[bsw::kDel + bsw::kGate + 2] != 1) {
    check_prep(h);  // the gate closed on a preparation error: nothing was written
    throw Error(-4, "delete batch: the update gate closed without an error word");
  }
  if (!fused) check_prep(h);
  stats_.last_ins_keys = stats_.last_ins_sources = 0;
  stats_.last_del_keys = h[2];
  stats_.last_del_sources = h[3];
  stats_.sum_del_sources += h[3];
  if (fused) {
    mutate([&] { delete_finish(h, k2_.get(), s32_[0].get(), dsd); });
    return;
  }
  if (h[2] == 0) return;
  // run_beg[nruns] = nR is written; the counts stay in dsc_[2], [3]
  delete_runs(k2_.get(), h[2], static_cast<uint32_t>(h[3]), dsc_.get() + 2);
}

// Delete sorted unique keys (absent keys are ignored).
void Graph::delete_sorted(const uint64_t *R0, uint64_t nR0) {
  if (nR0 == 0) return;
  if (read_only_) throw Error(-1, "graph is a query-only snapshot");
  GView g = view();
  // keep the keys that exist (absent deletes are no-ops)
  k2_.ensure(nR0 + 1, stream_);
  f0_.ensure(nR0 + 1, stream_);
  k_exists<<<grid_for(nR0, 256), 256, 0, stream_>>>(g, R0, nR0, f0_.get());  // one search per thread
  TG_CHECK_LAUNCH();
  select(SelFlagged{R0, f0_.get(), k2_.get()}, nR0, dsc_.get() + 2, scan_scratch, stream_);
  uint64_t h[kScalars];
  read_scalars(dsc_.get(), h, kScalars, stream_);
  check_prep(h);
  const uint64_t nR = h[2];
  if (nR == 0) return;
  uint64_t *R = k2_.get();
  DevBuf<uint32_t> &run_beg = s32_[0];
  run_beg.ensure(nR + 1, stream_);
  select(SelRunStart{R, run_beg.get()}, nR, dsc_.get() + 3, scan_scratch, stream_);
  k_put_u32<<<1, 1, 0, stream_>>>(run_beg.get(), dsc_.get() + 3, static_cast<uint32_t>(nR));
  TG_CHECK_LAUNCH();
  read_scalars(dsc_.get(), h, kScalars, stream_);
  check_prep(h);
  delete_runs(R, nR, static_cast<uint32_t>(h[3]), dsc_.get() + 2);  // counts in dsc_[2], [3]
}

// Two counts written to the device (bsw::kPut) for a half launched on
// counts the host knows.
const uint64_t *Graph::put_counts(uint64_t keys, uint64_t runs) {
  fill_u64(dsc_.get() + bsw::kPut, keys, 1, stream_);
  fill_u64(dsc_.get() + bsw::kPut + 1, runs, 1, stream_);
  return dsc_.get() + bsw::kPut;
}

// A mixed batch: the inserts, then the deletes (the driver's add_edge_batch
// followed by its remove_edge loop, src/graph_bm.cc:354-378), with one sort
// and one preparation for both: each edge's last op decides (an edge both
// inserted and deleted ends absent), deletes of absent edges are no-ops.
// The delete half's existence check runs before the insert half changes the
// graph: the two halves hold disjoint edges, so inserting one cannot change
// whether an edge of the other exists.
//
// Up to kBucketSortMax edges the whole batch is launched before the host
// reads anything: the preparation, the insert half (gated on the
// preparation's error words) and the delete half (gated also on the insert
// half completing on the device), then one read.  When the insert half
// needs a general pass (a promotion, a tree block or PMA leaf that does not
// fit in place), the delete half has written nothing; it runs again after
// that pass.
void Graph::update_batch(const uint32_t *is, const uint32_t *id, uint64_t ni, const uint32_t *ds,
                         const uint32_t *dd, uint64_t nd) {
  const uint64_t count = ni + nd;
  if (count == 0) return;
  TrimAfter trim_after{*this, count > kKeepScratch};
  if (ws_) throw Error(-1, "weighted graph: updates without weights");
  if (count > 0x7FFFFFFFull) throw Error(-1, "mixed batch larger than 2^31 edges");
  if (read_only_) throw Error(-1, "graph is a query-only snapshot");
  const bool small = count <= kSmallSortKeys;
  const bool fused = count <= kBucketSortMax;
  uint64_t *dsd = dsc_.get() + bsw::kDel;
  uint64_t h[kScalars];
  bool pma_launched = false;
  // deleted edges: fused, the unique keys D (kD_, count dsc_[9]) with their
  // existence flags (fD_), the prefix of the absent ones (s32_[2]) and their
  // source runs (s32_[1], count dsc_[10]); otherwise the existing ones
  // compacted into kE_ (count dsc_[10]), their runs in s32_[1] (dsc_[11])
  uint32_t *dnidx = nullptr;
  uint32_t *ileaf = nullptr;  // the inserted keys' PMA leaves from the preparation (bucket path)
  uint64_t *dslot = nullptr;  // the deleted keys' PMA slots from the preparation (fused paths)
  for (int pass = 0; pass < 2; pass++) {
    k1_.ensure(count + 1, stream_);
    kD_.ensure(count + 1, stream_);
    const bool bucket = small ? false : prepare_mixed_async(is, id, ni, ds, dd, nd, pass == 1);
    ileaf = nullptr;
    dslot = nullptr;
    GView g = view();
    const uint64_t *I = k1_.get(), *D = kD_.get();
    f0_.ensure(count + 1, stream_);
    f1_.ensure(count + 1, stream_);
    fD_.ensure(count + 1, stream_);
    s32_[0].ensure(count + 1, stream_);
    s32_[1].ensure(count + 1, stream_);
    s32_[2].ensure(count + 1, stream_);
    s32_[3].ensure(count + 1, stream_);
    dnidx = s32_[2].get();
    // inserted edges: existence, new-key prefix, source runs (run_beg in
    // s32_[0], count in dsc_[2])
    if (small) {
      // sort and split in one workgroup (also clears the batch scalars); the
      // rest of the preparation as for the bucket path
      k_small_mixed_sort<<<1, 1024, 0, stream_>>>(is, id, static_cast<uint32_t>(ni), ds, dd,
                                                  static_cast<uint32_t>(nd), n_, k1_.get(), kD_.get(), dsc_.get());
      TG_CHECK_LAUNCH();
    }
    if (count <= kBucketSortMax) {
      const uint32_t nb = static_cast<uint32_t>((count + kPrepBlock - 1) / kPrepBlock);
      const uint32_t epoch = scan_scratch.lb.begin(4ull * nb, stream_);
      ileaf = s32_[3].get();
      dslot_.ensure(count + 1, stream_);
      dslot = dslot_.get();
      k_mixed_prep<<<2 * nb, kPrepBlock, 0, stream_>>>(
          g, I, dsc_.get() + 1, count, f0_.get(), f1_.get(), s32_[0].get(), dsc_.get() + 2, ileaf, D, dsc_.get() + 9,
          fD_.get(), dnidx, s32_[1].get(), dsc_.get() + 10, dslot, scan_scratch.lb.state.get(),
          scan_scratch.lb.misc.get(), scan_scratch.lb.err_word(), nb, epoch);
      TG_CHECK_LAUNCH();
    } else {
      kE_.ensure(count + 1, stream_);
      k_exists2_n<<<grid_for(count, 256), 256, 0, stream_>>>(g, I, dsc_.get() + 1, count, f0_.get(), f1_.get());
      TG_CHECK_LAUNCH();
      const uint32_t nb = static_cast<uint32_t>((count + 8 * kPrepBlock - 1) / (8 * kPrepBlock));
      const uint32_t epoch = scan_scratch.lb.begin(2ull * nb, stream_);
      k_insert_prep<8, false><<<nb, kPrepBlock, 0, stream_>>>(
          g, I, dsc_.get() + 1, count, f0_.get(), f1_.get(), s32_[0].get(), dsc_.get() + 2,
          scan_scratch.lb.state.get(), scan_scratch.lb.misc.get(), scan_scratch.lb.err_word(), nb, epoch);
      TG_CHECK_LAUNCH();
      k_exists_n<<<grid_for(count, 256), 256, 0, stream_>>>(g, D, dsc_.get() + 9, count, fD_.get());
      TG_CHECK_LAUNCH();
      select(SelFlagged{D, fD_.get(), kE_.get()}, count, dsc_.get() + 10, scan_scratch, stream_);
      select(SelRunStartN{kE_.get(), s32_[1].get(), dsc_.get() + 10}, count, dsc_.get() + 11, scan_scratch, stream_);
    }
    if (debug_fail & 1u) {
      debug_fail &= ~1u;
      fill_u64(dsc_.get() + kErrScan, 1, 1, stream_);
    }
    if (fused) {
      // both halves on the device's counts (inserted keys and runs in
      // dsc_[1], [2]; deleted keys and runs in dsc_[9], [10]), their scratch
      // sized by the batch before the first write
      presize_insert(count, static_cast<uint32_t>(count));
      presize_delete(count, static_cast<uint32_t>(count));
      mutate([&] {
        pma_launched =
            insert_launch(k1_.get(), count, static_cast<uint32_t>(count), dsc_.get() + 1, dsc_.get(), ileaf);
        delete_launch(kD_.get(), fD_.get(), dnidx, count, static_cast<uint32_t>(count), s32_[1].get(),
                      dsc_.get() + 9, dsc_.get(), true, pma_launched, dsd, dslot);
      });
    }
    read_scalars(dsc_.get(), h, kScalars, stream_);
    if (!bucket || !(h[0] & kSortOverflow)) break;
  }
  if (fused && h[bsw::kGate + 2] != 1) {
    // the gate closed on a preparation error: nothing was written
    check_prep(h);
    throw Error(-4, "mixed batch: the update gate closed without an error word");
  }
  if (!fused) check_prep(h);
  const uint64_t nI = h[1], nIruns = h[2];
  stats_.last_ins_keys = nI;
  stats_.last_ins_sources = nIruns;
  stats_.sum_ins_sources += nIruns;
  if (!fused) {
    const uint64_t nE = h[10], nEruns = h[11];
    stats_.last_del_keys = nE;
    stats_.last_del_sources = nEruns;
    stats_.sum_del_sources += nEruns;
    if (nI) insert_runs(k1_.get(), nI, static_cast<uint32_t>(nIruns), dsc_.get() + 1);
    // the insert's general passes may have used dsc_[10], [11]
    if (nE) delete_runs(kE_.get(), nE, static_cast<uint32_t>(nEruns), put_counts(nE, nEruns), s32_[1].get());
    return;
  }
  const uint64_t nD = h[9], nDruns = h[10];
  stats_.last_del_sources = nDruns;  // sources of the deleted keys
  stats_.sum_del_sources += nDruns;
  mutate([&] {
    const bool ins_done = insert_finish(h, k1_.get(), pma_launched);
    const bool del_ran = h[bsw::kDel + bsw::kGate + 2] == 1;
    if (ins_done != del_ran) {
      check_flags(h);  // a look-back wait given up in the insert half closes the delete's gate
      throw Error(-4, "mixed batch: the delete half's gate disagrees with the insert half's outcome");
    }
    if (del_ran) {
      stats_.last_del_keys = h[bsw::kDel + 3];
      delete_finish(h, kD_.get(), s32_[1].get(), dsd);
    } else if (nD) {
      // the insert's general passes are done (they may have used dsc_[9],
      // [10]); the delete half now
      delete_launch(kD_.get(), fD_.get(), dnidx, nD, static_cast<uint32_t>(nDruns), s32_[1].get(),
                    put_counts(nD, nDruns), nullptr, false, false, dsd, dslot);
      uint64_t hs[kScalars];
      read_scalars(dsc_.get(), hs, kScalars, stream_);
      stats_.last_del_keys = hs[bsw::kDel + 3];
      delete_finish(hs, kD_.get(), s32_[1].get(), dsd);
    } else {
      stats_.last_del_keys = 0;
    }
  });
}

// Scratch of the delete's common path (see presize_insert).  Inline
// refills number at most the deleted keys.
void Graph::presize_delete(uint64_t nR, uint32_t nruns) {
  s32_[12].ensure(nruns + 1, stream_);
  s32_[13].ensure(nruns + 1, stream_);
  k3_.ensure(nR + 1, stream_);
  g0_.ensure(nR + 1, stream_);
  s64_[3].ensure(nR + 1, stream_);
  dhint_.ensure(nR + 1, stream_);
  s32_[30].ensure(1 + 3 * std::min<uint64_t>(nR, kTreeSmallKeys), stream_);
}

void Graph::delete_runs(const uint64_t *R, uint64_t nR, uint32_t nruns, const uint64_t *cw, const uint32_t *run_beg) {
  presize_delete(nR, nruns);
  const uint32_t *rb = run_beg ? run_beg : s32_[0].get();
  uint64_t *dsd = dsc_.get() + bsw::kDel;
  mutate([&] {
    delete_launch(R, nullptr, nullptr, nR, nruns, rb, cw, nullptr, false, false, dsd);
    uint64_t hs[kScalars];
    read_scalars(dsc_.get(), hs, kScalars, stream_);
    delete_finish(hs, R, rb, dsd);
  });
}

// The inline refill from the smallest secondary neighbours (k_bump; it
// removes the PMA elements it moves itself and counts them in ds[9]; the
// tree keys it moves go to s64_[3] at the offsets k_delete_runs gave, their
// count in ds[17]).
void Graph::delete_refill(const uint64_t *R, uint32_t capRuns, const uint32_t *run_beg, uint64_t *ds,
                          bool trees_only) {
  GView g = view();
  k_bump<<<grid_for(capRuns, 256), 256, 0, stream_>>>(g, R, run_beg, ds + bsw::kGate, s32_[12].get(), s32_[13].get(),
                                                      s64_[3].get(), trees_only, ds + 9,
                                                      static_cast<uint32_t>(bounds_.lo[0]), dsc_.get() + kLowFlag);
  TG_CHECK_LAUNCH();
  tree_dirty_ = true;
}

// The delete proper: R holds existing keys only, run_beg their run starts
// with run_beg[nruns] = nR, the counts on the device at cw.  Every stage
// runs on the device's counts (PMA deletes ds[4], tree deletes ds[5],
// inline refills ds[6]) with no host read between them: the PMA deletes,
// the in-place tree delete (status ds[13]), the inline refill, its tier
// split and the refill's in-place tree delete (status ds[19]), which waits
// when the first tree delete did not fit.
void Graph::delete_launch(const uint64_t *R, const uint32_t *exf, const uint32_t *dnidx, uint64_t capR,
                          uint32_t capRuns, const uint32_t *run_beg, const uint64_t *cw, const uint64_t *gate, bool spec,
                          bool pma_launched, uint64_t *ds, const uint64_t *dslot) {
  lists_dirty_ = true;
  ++epoch_;
  GView g = view();
  s32_[12].ensure(capRuns + 1, stream_);  // inline refills per run
  s32_[13].ensure(capRuns + 1, stream_);  // B-tree runs: offset of their refill keys
  k3_.ensure(capR + 1, stream_);
  g0_.ensure(capR + 1, stream_);
  s64_[3].ensure(capR + 1, stream_);
  if (dslot) dhint_.ensure(capR + 1, stream_);
  // inline deletes of every run and the secondary deletes split by tier
  // (PMA keys into k3_, count ds[4]; tree keys into g0_, ds[5])
  const uint32_t nb = static_cast<uint32_t>(std::max<uint64_t>(1, (capRuns + kRunBlock - 1) / kRunBlock));
  const uint32_t epoch = scan_scratch.lb.begin(4ull * nb, stream_);
  k_delete_runs<<<nb, kRunBlock, 0, stream_>>>(g, R, exf, dnidx, run_beg, cw, gate, spec, pma_launched, ds + bsw::kGate, capR,
                                               capRuns, k3_.get(), g0_.get(), s32_[12].get(), s32_[13].get(), ds,
                                               dsc_.get() + kLowFlag, dslot, dslot ? dhint_.get() : nullptr,
                                               scan_scratch.lb.state.get(),
                                               scan_scratch.lb.misc.get(), scan_scratch.lb.err_word(), nb, epoch);
  TG_CHECK_LAUNCH();
  if (debug_fail & 2u) {
    debug_fail &= ~2u;
    throw Error(-2, "injected allocation failure after the inline deletes");
  }
  // the tree delete beside the PMA delete for a large batch (the refill
  // reads both tiers after the join)
  const bool side = capR >= kSideMin;
  if (side) fork_side(2);
  tree_small_launch(g0_.get(), capR, false, ds + 5, nullptr, ds + 13, side ? side_ : nullptr);
  pma_delete_launch(k3_.get(), capR, ds + 4, dslot ? dhint_.get() : nullptr);
  if (side) join_side(3);
  delete_refill(R, capRuns, run_beg, ds, false);
  tree_small_launch(s64_[3].get(), capR, false, ds + 17, ds + 13, ds + 19);
}

void Graph::delete_finish(const uint64_t *hs, const uint64_t *R, const uint32_t *run_beg, uint64_t *ds) {
  const uint64_t *hd = hs + (ds - dsc_.get());
  const uint64_t nR = hd[bsw::kGate];
  const uint32_t nruns = static_cast<uint32_t>(hd[bsw::kGate + 1]);
  const uint64_t npma = hd[4], ntree = hd[5], nb = hd[6];
  DevBuf<uint64_t> &btree = s64_[3];
  pma_elems_ -= npma + hd[9];  // deleted PMA keys, PMA elements the refill moved inline
  bool fallback = false;
  if (ntree && hd[13] == 0) stats_.tree_small_updates++;
  if (ntree && hd[13] != 0) {
    // the in-place tree delete did not fit: copy-on-write, then the refill
    // of the B-tree vertices again (its writes are positional, so running
    // it again gives the same result) and the delete of its tree keys
    fallback = true;
    tree_delete_cow(g0_.get(), ntree);
    if (nb) {
      fill_u64(ds + 9, 0, 1, stream_);  // a trees-only refill moves no PMA element
      delete_refill(R, nruns, run_beg, ds, true);
      tree_delete(btree.get(), hd[17]);  // the same count: B-tree runs refill as before
    }
  } else if (hd[17]) {
    if (hd[19] == 0) {
      stats_.tree_small_updates++;
    } else {
      fallback = true;
      tree_delete_cow(btree.get(), hd[17]);
    }
  }
  if (ws_) {
    ws_->erase(R, nR, stream_);  // R holds exactly the deleted (existing) keys
    check_batch_flags();
  } else if (fallback) {
    check_batch_flags();
  } else {
    check_flags(hs);
  }
  // the flag word is only set after k_delete_runs cleared it: the fallback's
  // deletes add to it
  if (fallback ? read_scalar(dsc_.get() + kLowFlag, stream_) != 0 : hs[kLowFlag] != 0) pma_rebalance_low();
}

// ------------------------------------------------------------ weights
void Graph::enable_weights() {
  if (ws_) return;
  if (num_edges() != 0) throw Error(-1, "weights must be enabled on an empty graph");
  ws_ = std::make_unique<WeightStore>();
}

void Graph::insert_batch_weighted(const uint32_t *src, const uint32_t *dst, const uint32_t *w, uint64_t count) {
  if (count == 0) return;
  TrimAfter trim_after{*this, count > kKeepScratch};
  if (!ws_) throw Error(-1, "graph is not weighted");
  if (count > 0xFFFFFFFFull) throw Error(-1, "batch larger than 2^32 edges");
  if (read_only_) throw Error(-1, "graph is a query-only snapshot");
  k0_.ensure(count, stream_);
  k1_.ensure(count + 1, stream_);
  wv0_.ensure(count, stream_);
  wv1_.ensure(count + 1, stream_);
  fill_u64(dsc_.get(), 0, kErrSort, stream_);
  k_pack<<<grid_for(count, 256), 256, 0, stream_>>>(src, dst, count, n_, k0_.get(), dsc_.get());
  TG_CHECK_LAUNCH();
  k_check_weights<<<grid_for(count, 256), 256, 0, stream_>>>(w, count, dsc_.get());
  TG_CHECK_LAUNCH();
  copy_bytes(wv0_.get(), w, count * 4, stream_);
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < n_) bits++;
  radix_sort_pairs(k0_.get(), wv0_.get(), count, bits, bits, sort_, stream_);  // stable: input order within a key
  select(SelUniqueW{k0_.get(), wv0_.get(), k1_.get(), wv1_.get()}, count, dsc_.get() + 1, scan_scratch, stream_);
  uint64_t h[kScalars];
  read_scalars(dsc_.get(), h, kScalars, stream_);
  if (h[0] & 8) {
    fill_u64(dsc_.get(), 0, 1, stream_);
    throw Error(-1, "edge weight must be below 2^31");
  }
  check_prep(h);
  uint64_t nR = h[1];
  // the keys absent before the insert get their weights
  wput_.ensure(nR + 1, stream_);
  s32_[29].ensure(nR + 1, stream_);
  k_exists2<<<grid_for(nR, 256), 256, 0, stream_>>>(view(), k1_.get(), nR, s32_[29].get(), wput_.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u32(wput_.get(), s32_[29].get(), nR, s32_[29].get() + nR, scan_scratch, stream_);
  uint64_t nnew = read_scalar(s32_[29].get() + nR, stream_);
  insert_sorted(k1_.get(), nR);
  mutate([&] {
    ws_->insert(k1_.get(), wv1_.get(), wput_.get(), nR, nnew, stream_);
    TG_HIP(hipStreamSynchronize(stream_));
  });
}

void Graph::edge_weights(const uint32_t *src, const uint32_t *dst, uint64_t count, uint32_t *out) {
  if (!ws_) throw Error(-1, "graph is not weighted");
  ws_->lookup(src, dst, count, out, stream_);
}

// ------------------------------------------------------------ coalesced ops
// A coalesced run of update batches: ops[i] = 0 insert / 1 delete, in
// arrival order.  The stable sort keeps arrival order inside each key, so
// the last op of a key decides, which equals applying the batches one by
// one (both ops are idempotent set updates).
void Graph::apply_ops(const uint32_t *src, const uint32_t *dst, const uint32_t *ops, uint64_t count) {
  if (count == 0) return;
  TrimAfter trim_after{*this, count > kKeepScratch};
  if (ws_) throw Error(-1, "weighted graph: coalesced ops carry no weights");
  if (count > 0xFFFFFFFFull) throw Error(-1, "batch larger than 2^32 edges");
  k0_.ensure(count, stream_);
  fill_u64(dsc_.get(), 0, kErrSort, stream_);  // the look-back error words stay until checked
  k_pack<<<grid_for(count, 256), 256, 0, stream_>>>(src, dst, count, n_, k0_.get(), dsc_.get());
  TG_CHECK_LAUNCH();
  ops_val_.ensure(count, stream_);
  TG_HIP(hipMemcpyAsync(ops_val_.get(), ops, count * 4, hipMemcpyDeviceToDevice, stream_));
  uint32_t bits = 1;
  while (bits < 32 && (1ull << bits) < n_) bits++;
  radix_sort_pairs(k0_.get(), ops_val_.get(), count, bits, bits, sort_, stream_);
  ops_ins_.ensure(count + 1, stream_);
  ops_del_.ensure(count + 1, stream_);
  select(SelLastOp{k0_.get(), ops_val_.get(), count, 0, ops_ins_.get()}, count, dsc_.get() + 1, scan_scratch,
         stream_);
  select(SelLastOp{k0_.get(), ops_val_.get(), count, 1, ops_del_.get()}, count, dsc_.get() + 2, scan_scratch,
         stream_);
  uint64_t h[kScalars];
  read_scalars(dsc_.get(), h, kScalars, stream_);
  check_prep(h);
  insert_sorted(ops_ins_.get(), h[1]);
  delete_sorted(ops_del_.get(), h[2]);
}

// ------------------------------------------------------------ queries
void Graph::degrees(uint32_t *out) const {
  k_degrees<<<grid_for(n_, 256), 256, 0, stream_>>>(view(), out);
  TG_CHECK_LAUNCH();
}

uint64_t Graph::num_edges() {
  g1_.ensure(static_cast<uint64_t>(n_) + 1, stream_);
  k_degrees_u64<<<grid_for(n_, 256), 256, 0, stream_>>>(view(), g1_.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u64(g1_.get(), g1_.get(), n_, g1_.get() + n_, scan_scratch, stream_);
  return read_scalar(g1_.get() + n_, stream_);
}

void Graph::count_common(const uint32_t *src, const uint32_t *dst, uint64_t q, uint32_t *out, uint64_t *err) {
  if (!q) return;
  GView g = view();
  if (!err) {
    cc_err_.ensure(1, stream_);
    err = cc_err_.get();
  }
  fill_u64(err, 0, 1, stream_);  // a kernel, no memset node
  k_count_common<<<grid_for(q * 64, 256, 1u << 16), 256, 0, stream_>>>(g, src, dst, q, out, err);
  TG_CHECK_LAUNCH();
  if (read_scalar(err, stream_)) throw Error(-1, "count_common: vertex id out of range");
}

uint64_t Graph::count_common_edges(const uint32_t *list, uint64_t cnt, unsigned long long *total) {
  GView g = view();
  const uint64_t rows = list ? cnt : n_;
  cc_off_.ensure(static_cast<uint64_t>(n_) + 1, stream_);
  const uint64_t m = num_edges();
  cc_nbr_.ensure(m + 1, stream_);
  export_adjacency(cc_off_.get(), cc_nbr_.get());
  if (rows == 0) return 0;
  DevBuf<uint32_t> lower;
  DevBuf<uint64_t> pos;
  lower.ensure(rows + 1, stream_);
  pos.ensure(rows + 1, stream_);
  k_tc_lower<<<grid_for(rows, 256), 256, 0, stream_>>>(cc_off_.get(), cc_nbr_.get(), list, rows, lower.get());
  TG_CHECK_LAUNCH();
  k_u32_to_u64_n<<<grid_for(rows, 256), 256, 0, stream_>>>(lower.get(), rows, pos.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u64(pos.get(), pos.get(), rows, pos.get() + rows, scan_scratch, stream_);
  const uint64_t np = read_scalar(pos.get() + rows, stream_);
  if (np == 0) return 0;
  DevBuf<uint64_t> pairs;
  pairs.ensure(np, stream_);
  k_tc_pairs<<<grid_for(rows * 64, 256, 1u << 16), 256, 0, stream_>>>(cc_off_.get(), cc_nbr_.get(), list, rows,
                                                                       pos.get(), pairs.get());
  TG_CHECK_LAUNCH();
  k_tc_count_pairs<<<grid_for(np * 64, 256, 1u << 16), 256, 0, stream_>>>(g, pairs.get(), np, total);
  TG_CHECK_LAUNCH();
  TG_HIP(hipStreamSynchronize(stream_));
  return np;
}

void Graph::is_edge_batch(const uint32_t *src, const uint32_t *dst, uint64_t count, uint8_t *out) {
  if (!count) return;
  k_is_edge<<<grid_for(count, 256), 256, 0, stream_>>>(view(), src, dst, count, out);
  TG_CHECK_LAUNCH();
}

void Graph::tiers(uint8_t *out) const {
  k_tiers<<<grid_for(n_, 256), 256, 0, stream_>>>(view(), out);
  TG_CHECK_LAUNCH();
}

void Graph::tier_counts(uint64_t out[4]) {
  g1_.ensure(4, stream_);
  fill_u64(g1_.get(), 0, 4, stream_);
  k_tier_counts<<<grid_for(n_, 256, 4096), 256, 0, stream_>>>(view(), reinterpret_cast<unsigned long long *>(g1_.get()));
  TG_CHECK_LAUNCH();
  read_scalars(g1_.get(), out, 4, stream_);
}

const Graph::TierLists &Graph::tier_lists() {
  if (!lists_dirty_) return lists_;
  // the two lists by predicate-fused selects over the vertex blocks (sized
  // by n), both counts in one read
  list_pma_.ensure(static_cast<uint64_t>(n_) + 1, stream_);
  list_tree_.ensure(static_cast<uint64_t>(n_) + 1, stream_);
  select2(SelTiers{vb_.get(), list_pma_.get(), list_tree_.get()}, n_, dsc_.get() + 11, dsc_.get() + 12, scan_scratch,
          stream_);
  uint64_t h[2];
  read_scalars(dsc_.get() + 11, h, 2, stream_);
  const uint32_t np = static_cast<uint32_t>(h[0]), nt = static_cast<uint32_t>(h[1]);
  lists_.pma = list_pma_.get();
  lists_.npma = np;
  lists_.tree = list_tree_.get();
  lists_.ntree = nt;
  lists_dirty_ = false;
  return lists_;
}

void Graph::export_adjacency(uint64_t *offsets, uint32_t *nbrs) {
  GView g = view();
  k_degrees_u64<<<grid_for(n_, 256), 256, 0, stream_>>>(g, offsets);
  TG_CHECK_LAUNCH();
  exclusive_scan_u64(offsets, offsets, n_, offsets + n_, scan_scratch, stream_);
  const TierLists &L = tier_lists();
  k_export_inline<<<grid_for(static_cast<uint64_t>(n_) * 16, 256), 256, 0, stream_>>>(g, offsets, nbrs);
  TG_CHECK_LAUNCH();
  if (L.npma) {
    k_export_pma<<<grid_for(static_cast<uint64_t>(L.npma) * 16, 256), 256, 0, stream_>>>(g, L.pma, L.npma, offsets,
                                                                                         nbrs);
    TG_CHECK_LAUNCH();
  }
  if (L.ntree) {
    k_export_tree<<<grid_for(static_cast<uint64_t>(L.ntree) * 256, 256, 1u << 16), 256, 0, stream_>>>(
        g, L.tree, L.ntree, offsets, nbrs);
    TG_CHECK_LAUNCH();
  }
  TG_HIP(hipStreamSynchronize(stream_));
}

void Graph::neighbors_of(uint32_t v, uint32_t *out) {
  if (v >= n_) throw Error(-1, "vertex out of range");
  k_neighbors_one<<<1, 64, 0, stream_>>>(view(), v, out);
  TG_CHECK_LAUNCH();
  TG_HIP(hipStreamSynchronize(stream_));
}

void Graph::copy_from(Graph &src, bool query_only) {
  if (src.n_ != n_) throw Error(-1, "snapshot of a graph with another vertex count");
  if (src.invalid_) throw Error(-4, "snapshot of an invalid graph");
  src.rebuild_count_tree();
  hipStream_t s = src.stream_;
  auto cp = [&](void *d, const void *p, uint64_t bytes) {
    if (bytes) TG_HIP(hipMemcpyAsync(d, p, bytes, hipMemcpyDeviceToDevice, s));
  };
  TG_HIP(hipStreamSynchronize(stream_));
  cp(vb_.get(), src.vb_.get(), sizeof(VBlock) * static_cast<uint64_t>(n_));
  if (pma_N_ != src.pma_N_) {
    // a query-only copy keeps a slot array that is already large enough
    // (repeated snapshots of a growing graph then do not reallocate); an
    // updatable copy starts its update-side state afresh
    if (!query_only || pdst_.cap < src.pma_N_) pdst_.release();
    psrc_.release();
    pcnt_.release();
    fence_.release();
    mark_.release();
    pma_H_ = src.pma_H_;
    pma_N_ = src.pma_N_;
    pdst_.ensure(pma_N_, s);
    bounds_ = src.bounds_;
  }
  uint64_t nl = 1ull << pma_H_;
  cp(pdst_.get(), src.pdst_.get(), pma_N_ * 4);
  if (!query_only) {
    // update-side PMA state (a query-only snapshot never allocates it)
    psrc_.ensure(pma_N_, s);
    pcnt_.ensure(2 * nl, s);
    fence_.ensure(nl, s);
    if (mark_.ensure(2 * nl, s)) TG_HIP(hipMemsetAsync(mark_.get(), 0, 2 * nl, s));
    cp(psrc_.get(), src.psrc_.get(), pma_N_ * 4);
    cp(pcnt_.get(), src.pcnt_.get(), 2 * nl * 4);
    cp(fence_.get(), src.fence_.get(), nl * 8);
  }
  tree_dirty_ = false;
  read_only_ = query_only;
  tune = src.tune;
  ensure_blocks(src.blk_top_);
  ensure_index(src.idx_top_);
  ensure_trees(src.ntrees_ + 1);
  cp(bkeys_.get(), src.bkeys_.get(), src.blk_top_ * kBlk * 4);
  cp(bcnt_.get(), src.bcnt_.get(), src.blk_top_ * 4);
  cp(iblk_.get(), src.iblk_.get(), src.idx_top_ * 4);
  cp(imin_.get(), src.imin_.get(), src.idx_top_ * 4);
  cp(toff_.get(), src.toff_.get(), src.ntrees_ * 8ull);
  cp(tnblk_.get(), src.tnblk_.get(), src.ntrees_ * 4ull);
  cp(tvtx_.get(), src.tvtx_.get(), src.ntrees_ * 4ull);
  blk_top_ = src.blk_top_;
  idx_top_ = src.idx_top_;
  pma_elems_ = src.pma_elems_;
  ntrees_ = src.ntrees_;
  lists_dirty_ = true;
  ++epoch_;
  TG_HIP(hipStreamSynchronize(s));
}

uint64_t Graph::memory_bytes() const {
  return (ws_ ? ws_->bytes() : 0) + vb_.cap * sizeof(VBlock) + (pdst_.cap + psrc_.cap + pcnt_.cap) * 4 + fence_.cap * 8 + mark_.cap +
         (bkeys_.cap + bcnt_.cap + iblk_.cap + imin_.cap) * 4 + toff_.cap * 8 + (tnblk_.cap + tvtx_.cap) * 4;
}

}  // namespace tg
