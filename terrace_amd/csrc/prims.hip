// Device primitives for gfx950: reduce-then-scan exclusive scan and a
// stable LSD radix sort of 64-bit (src,dst) keys with wave64 multisplit
// ranking.  The sort replaces the reference's integerSort_y/integerSort_x
// pair (include/integerSort/blockRadixSort/blockRadixSort.h:287-292), whose
// output is the lexicographic (src,dst) order produced here.
#include "prims.hpp"

namespace tg {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;  // 2048

template <typename T>
__device__ T wave_incl_scan(T v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T o = __shfl_up(v, off, 64);
    if (lane >= static_cast<uint32_t>(off)) v += o;
  }
  return v;
}

template <typename T>
__device__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Exclusive scan of the 256 per-thread values of a block; returns the
// thread's exclusive prefix and writes the block total into *total.
template <typename T>
__device__ T block_excl_scan(T v, T *lds_waves, T *total) {
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  T inc = wave_incl_scan(v);
  if ((tid & 63) == 63) lds_waves[w] = inc;
  __syncthreads();
  T wave_off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanBlock / 64; i++) {
    T x = lds_waves[i];
    if (i < w) wave_off += x;
    tot += x;
  }
  *total = tot;
  __syncthreads();
  return wave_off + inc - v;
}

template <typename InT, typename AccT>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const InT *in, uint64_t n, AccT *sums) {
  __shared__ AccT lds[kScanBlock / 64];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  AccT acc = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kScanBlock + threadIdx.x;
    if (idx < n) acc += static_cast<AccT>(in[idx]);
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    AccT t = 0;
    for (int i = 0; i < kScanBlock / 64; i++) t += lds[i];
    sums[blockIdx.x] = t;
  }
}

// Block-level scan of a tile with per-block offsets[blockIdx.x] added.  The
// tile is staged through LDS so global loads and stores stay coalesced.
template <typename InT, typename AccT>
__global__ void __launch_bounds__(kScanBlock) k_scan_tile(const InT *in, InT *out, uint64_t n,
                                                          const AccT *offsets, InT *total_dev,
                                                          uint64_t *total64) {
  __shared__ AccT lds[kScanBlock / 64];
  __shared__ InT tile[kScanTile];
  const uint64_t tbase = static_cast<uint64_t>(blockIdx.x) * kScanTile;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = tbase + static_cast<uint64_t>(i) * kScanBlock + threadIdx.x;
    tile[i * kScanBlock + threadIdx.x] = idx < n ? in[idx] : InT(0);
  }
  __syncthreads();
  AccT vals[kScanItems];
  AccT sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    vals[i] = static_cast<AccT>(tile[threadIdx.x * kScanItems + i]);
    sum += vals[i];
  }
  AccT tot;
  AccT pre = block_excl_scan(sum, lds, &tot);
  AccT off = offsets ? offsets[blockIdx.x] : 0;
  pre += off;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    tile[threadIdx.x * kScanItems + i] = static_cast<InT>(pre);
    pre += vals[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = tbase + static_cast<uint64_t>(i) * kScanBlock + threadIdx.x;
    if (idx < n) out[idx] = tile[i * kScanBlock + threadIdx.x];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    if (total_dev) *total_dev = static_cast<InT>(off + tot);
    if (total64) *total64 = static_cast<uint64_t>(off + tot);
  }
}

// Single-pass scan: tile scan plus the decoupled look-back for the block's
// prefix (one launch for any n).
template <typename InT>
__global__ void __launch_bounds__(kScanBlock) k_scan_lb(const InT *in, InT *out, uint64_t n, InT *total_dev,
                                                        uint64_t *total64,
                                                        uint64_t *state, uint32_t *misc, uint32_t *err,
                                                        uint32_t nblocks, uint32_t epoch) {
  __shared__ uint64_t lds[kScanBlock / 64];
  __shared__ InT tile[kScanTile];
  __shared__ uint32_t s_ticket;
  __shared__ uint64_t s_excl;
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;  // inconsistent ticket counter: reported, no work
  const uint64_t tbase = static_cast<uint64_t>(b) * kScanTile;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = tbase + static_cast<uint64_t>(i) * kScanBlock + threadIdx.x;
    tile[i * kScanBlock + threadIdx.x] = idx < n ? in[idx] : InT(0);
  }
  __syncthreads();
  uint64_t vals[kScanItems];
  uint64_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    vals[i] = static_cast<uint64_t>(tile[threadIdx.x * kScanItems + i]);
    sum += vals[i];
  }
  uint64_t tot;
  uint64_t pre = block_excl_scan(sum, lds, &tot);
  if (threadIdx.x < 64) {  // wave 0 walks the look-back
    const uint32_t x = lbk::lookback_wave(state, 1, b, 0, err, epoch, static_cast<uint32_t>(tot));
    if (threadIdx.x == 0) s_excl = x;
  }
  __syncthreads();
  const uint64_t excl = s_excl;
  pre += excl;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    tile[threadIdx.x * kScanItems + i] = static_cast<InT>(pre);
    pre += vals[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    uint64_t idx = tbase + static_cast<uint64_t>(i) * kScanBlock + threadIdx.x;
    if (idx < n) out[idx] = tile[i * kScanBlock + threadIdx.x];
  }
  if (b == nblocks - 1 && threadIdx.x == 0) {
    if (total_dev) *total_dev = static_cast<InT>(excl + tot);
    if (total64) *total64 = excl + tot;
  }
}

template <typename T>
__global__ void k_set_zero(T *p) { *p = 0; }

template <typename InT>
void scan_impl(const InT *in, InT *out, uint64_t n, InT *total_dev, ScanScratch &ss, int level,
               hipStream_t s, uint64_t *total64 = nullptr) {
  if (n == 0) {
    if (total_dev) {
      k_set_zero<InT><<<1, 1, 0, s>>>(total_dev);
      TG_CHECK_LAUNCH();
    }
    if (total64) {
      k_set_zero<uint64_t><<<1, 1, 0, s>>>(total64);
      TG_CHECK_LAUNCH();
    }
    return;
  }
  uint64_t nblocks = (n + kScanTile - 1) / kScanTile;
  if (nblocks == 1) {
    k_scan_tile<InT, uint64_t><<<1, kScanBlock, 0, s>>>(in, out, n, nullptr, total_dev, total64);
    TG_CHECK_LAUNCH();
    return;
  }
  // single pass for 32-bit scans (prefixes fit the look-back word)
  if (ss.lookback && sizeof(InT) == 4 && nblocks <= 0xFFFFFFFFull) {
    uint32_t nb = static_cast<uint32_t>(nblocks);
    uint32_t epoch = ss.lb.begin(nb, s);
    k_scan_lb<InT><<<nb, kScanBlock, 0, s>>>(in, out, n, total_dev, total64, ss.lb.state.get(), ss.lb.misc.get(),
                                             ss.lb.err_word(), nb, epoch);
    TG_CHECK_LAUNCH();
    return;
  }
  if (level >= 5) throw Error(-1, "scan: input too large");
  DevBuf<uint64_t> &sums = ss.levels[level];
  sums.ensure(nblocks, s);
  k_scan_reduce<InT, uint64_t><<<static_cast<unsigned>(nblocks), kScanBlock, 0, s>>>(in, n, sums.get());
  TG_CHECK_LAUNCH();
  scan_impl<uint64_t>(sums.get(), sums.get(), nblocks, nullptr, ss, level + 1, s);
  k_scan_tile<InT, uint64_t><<<static_cast<unsigned>(nblocks), kScanBlock, 0, s>>>(in, out, n, sums.get(),
                                                                                  total_dev, total64);
  TG_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ sort
constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;  // 2048 keys
constexpr int kWaves = kSortBlock / 64;

// Key index of item j of a lane: waves own contiguous 512-key sub-tiles,
// iteration j covers 64 consecutive keys, so (wave, j, lane) is index order.
__device__ inline uint64_t sort_index(uint64_t tile_base, int w, int j, int lane) {
  return tile_base + static_cast<uint64_t>(w) * (kSortItems * 64) + j * 64 + lane;
}

__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const uint64_t *keys, uint64_t n, int shift,
                                                           uint32_t *hist, uint32_t nblocks) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kSortTile;
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    if (idx < n) atomicAdd(&cnt[(keys[idx] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  hist[static_cast<uint64_t>(threadIdx.x) * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const uint64_t *keys, uint64_t *out, uint64_t n,
                                                              int shift, const uint32_t *offs,
                                                              uint32_t nblocks, const uint32_t *vals,
                                                              uint32_t *vout) {
  __shared__ uint32_t wave_cnt[kWaves][256];
  __shared__ uint32_t digit_base[256];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < kWaves * 256; i += kSortBlock) (&wave_cnt[0][0])[i] = 0;
  __syncthreads();
  uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kSortTile;
  uint64_t k[kSortItems];
  uint32_t rank[kSortItems];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    bool valid = idx < n;
    k[j] = valid ? keys[idx] : 0;
    uint32_t d = static_cast<uint32_t>((k[j] >> shift) & 0xFF);
    // peers = lanes of this wave holding the same digit (valid only)
    uint64_t active = __ballot(valid);
    uint64_t peers = active;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    uint32_t r = __popcll(peers & lt);
    int leader = valid ? __ffsll(static_cast<unsigned long long>(peers)) - 1 : 0;
    uint32_t base = 0;
    if (valid && lane == leader) {
      base = wave_cnt[w][d];
      wave_cnt[w][d] = base + static_cast<uint32_t>(__popcll(peers));
    }
    base = __shfl(base, leader, 64);
    rank[j] = base + r;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;  // 256 threads == 256 digits
    uint32_t run = offs[static_cast<uint64_t>(d) * nblocks + blockIdx.x];
    digit_base[d] = run;
    for (int ww = 0; ww < kWaves; ww++) {
      uint32_t c = wave_cnt[ww][d];
      wave_cnt[ww][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    if (idx < n) {
      uint32_t d = static_cast<uint32_t>((k[j] >> shift) & 0xFF);
      uint32_t pos = wave_cnt[w][d] + rank[j];
      out[pos] = k[j];
      if (vals) vout[pos] = vals[idx];
    }
  }
  (void)digit_base;
}

// ---- onesweep radix sort: all digit histograms in one pass over the keys,
// then one kernel per digit pass whose blocks find their per-digit offsets
// by decoupled look-back (one packed word per (block, digit): epoch << 34 |
// status << 32 | count).
struct SortShifts {
  int s[8];
  int np;
};

__global__ void __launch_bounds__(kSortBlock) k_radix_hist_all(const uint64_t *keys, uint64_t n, SortShifts sh,
                                                               uint32_t *ghist) {
  __shared__ uint32_t cnt[8][256];
  for (int p = 0; p < sh.np; p++) cnt[p][threadIdx.x] = 0;
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t tile = static_cast<uint64_t>(blockIdx.x) * kSortTile;
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    if (idx < n) {
      uint64_t k = keys[idx];
      for (int p = 0; p < sh.np; p++) atomicAdd(&cnt[p][(k >> sh.s[p]) & 0xFF], 1u);
    }
  }
  __syncthreads();
  for (int p = 0; p < sh.np; p++) {
    uint32_t c = cnt[p][threadIdx.x];
    if (c) atomicAdd(&ghist[p * 256 + threadIdx.x], c);
  }
}

__global__ void __launch_bounds__(kSortBlock) k_onesweep(const uint64_t *keys, uint64_t *out, uint64_t n, int shift,
                                                         const uint32_t *ghist, const uint32_t *vals,
                                                         uint32_t *vout, uint64_t *state, uint32_t *misc,
                                                         uint32_t *err, uint32_t nblocks, uint32_t epoch) {
  __shared__ uint32_t wave_cnt[kWaves][256];
  __shared__ uint64_t lds[kSortBlock / 64];
  __shared__ uint32_t s_ticket;
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < kWaves * 256; i += kSortBlock) (&wave_cnt[0][0])[i] = 0;
  __syncthreads();
  uint64_t tile = static_cast<uint64_t>(b) * kSortTile;
  uint64_t k[kSortItems];
  uint32_t rank[kSortItems];
  int lead[kSortItems];
  const uint64_t lt = lanemask_lt();
  // all key loads issued first
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    k[j] = idx < n ? keys[idx] : 0;
  }
  // each group of equal digits takes its place with one LDS atomic by its
  // leader: a wave's atomics complete in issue order, so ranks follow the
  // index order without one step waiting on the one before
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    bool valid = idx < n;
    uint32_t d = static_cast<uint32_t>((k[j] >> shift) & 0xFF);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bb = 0; bb < 8; bb++) {
      uint64_t m = __ballot((d >> bb) & 1u);
      peers &= ((d >> bb) & 1u) ? m : ~m;
    }
    lead[j] = valid ? __ffsll(static_cast<unsigned long long>(peers)) - 1 : 0;
    uint32_t old = 0;
    if (valid && lane == static_cast<uint32_t>(lead[j]))
      old = atomicAdd(&wave_cnt[w][d], static_cast<uint32_t>(__popcll(peers)));
    rank[j] = old + static_cast<uint32_t>(__popcll(peers & lt));  // the leader's lt-count is 0
  }
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    const uint32_t r = rank[j];
    const uint32_t base = __shfl(r, lead[j], 64);
    rank[j] = lane == static_cast<uint32_t>(lead[j]) ? r : base + r;
  }
  __syncthreads();
  const uint32_t d = threadIdx.x;  // 256 threads == 256 digits
  uint32_t c = 0;
  for (int ww = 0; ww < kWaves; ww++) c += wave_cnt[ww][d];
  // global start of digit d: exclusive scan of the pass histogram
  uint64_t gtot;
  uint64_t gbase = block_excl_scan(static_cast<uint64_t>(ghist[d]), lds, &gtot);
  // block prefix of digit d over the blocks before b
  uint64_t excl = lbk::lookback_window(state, 256, b, d, err, epoch, c);
  uint64_t run = gbase + excl;
  for (int ww = 0; ww < kWaves; ww++) {
    uint32_t x = wave_cnt[ww][d];
    wave_cnt[ww][d] = static_cast<uint32_t>(run);
    run += x;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; j++) {
    uint64_t idx = sort_index(tile, w, j, lane);
    if (idx < n) {
      uint32_t dd = static_cast<uint32_t>((k[j] >> shift) & 0xFF);
      uint32_t pos = wave_cnt[w][dd] + rank[j];
      out[pos] = k[j];
      if (vals) vout[pos] = vals[idx];
    }
  }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

}  // namespace

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total_dev,
                        ScanScratch &ss, hipStream_t s, uint64_t *total64) {
  scan_impl<uint32_t>(in, out, n, total_dev, ss, 0, s, total64);
}

void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                        ScanScratch &ss, hipStream_t s) {
  scan_impl<uint64_t>(in, out, n, total_dev, ss, 0, s);
}

void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t n, uint32_t lo_bits, uint32_t hi_bits,
                      SortScratch &ss, hipStream_t s) {
  if (n < 2) return;
  if (n > 0xFFFFFFFFull) throw Error(-1, "radix sort: more than 2^32 keys");
  uint32_t nblocks = static_cast<uint32_t>((n + kSortTile - 1) / kSortTile);
  ss.alt.ensure(n, s);
  if (vals) ss.alt_vals.ensure(n, s);
  ss.hist.ensure(static_cast<uint64_t>(nblocks) * 256, s);
  uint64_t *src = keys, *dst = ss.alt.get();
  uint32_t *vsrc = vals, *vdst = vals ? ss.alt_vals.get() : nullptr;
  if (ss.scan.lookback) {
    SortShifts sh{};
    for (uint32_t b = 0; b < lo_bits; b += 8) sh.s[sh.np++] = static_cast<int>(b);
    for (uint32_t b = 0; b < hi_bits; b += 8) sh.s[sh.np++] = static_cast<int>(32 + b);
    if (sh.np > 8) throw Error(-1, "radix sort: more than 8 digit passes");
    ss.hist.ensure(8 * 256, s);
    fill_u32(ss.hist.get(), 0, static_cast<uint64_t>(sh.np) * 256, s);
    k_radix_hist_all<<<nblocks, kSortBlock, 0, s>>>(src, n, sh, ss.hist.get());
    TG_CHECK_LAUNCH();
    for (int p = 0; p < sh.np; p++) {
      uint32_t epoch = ss.scan.lb.begin(static_cast<uint64_t>(nblocks) * 256, s);
      k_onesweep<<<nblocks, kSortBlock, 0, s>>>(src, dst, n, sh.s[p], ss.hist.get() + p * 256, vsrc, vdst,
                                                ss.scan.lb.state.get(), ss.scan.lb.misc.get(), ss.scan.lb.err_word(),
                                                nblocks, epoch);
      TG_CHECK_LAUNCH();
      std::swap(src, dst);
      std::swap(vsrc, vdst);
    }
    if (src != keys) {
      TG_HIP(hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
      if (vals) TG_HIP(hipMemcpyAsync(vals, vsrc, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
    return;
  }
  auto pass = [&](int shift) {
    k_radix_hist<<<nblocks, kSortBlock, 0, s>>>(src, n, shift, ss.hist.get(), nblocks);
    TG_CHECK_LAUNCH();
    exclusive_scan_u32(ss.hist.get(), ss.hist.get(), static_cast<uint64_t>(nblocks) * 256, nullptr,
                       ss.scan, s);
    k_radix_scatter<<<nblocks, kSortBlock, 0, s>>>(src, dst, n, shift, ss.hist.get(), nblocks, vsrc, vdst);
    TG_CHECK_LAUNCH();
    std::swap(src, dst);
    std::swap(vsrc, vdst);
  };
  for (uint32_t b = 0; b < lo_bits; b += 8) pass(static_cast<int>(b));
  for (uint32_t b = 0; b < hi_bits; b += 8) pass(static_cast<int>(32 + b));
  if (src != keys) {
    TG_HIP(hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    if (vals) TG_HIP(hipMemcpyAsync(vals, vsrc, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
}

namespace {
constexpr uint32_t kSortSmallKeys = 4096;
// At most kSortSmallKeys keys: one workgroup sorts them in LDS (bitonic,
// padded with ~0 to a power of two) and writes them back in place.  Keys
// carry no values, so the order equals the radix sort's.
__global__ void __launch_bounds__(1024) k_sort_small_u64(uint64_t *keys, uint32_t n) {
  __shared__ uint64_t sk[kSortSmallKeys];
  const uint32_t t = threadIdx.x;
  uint32_t P = 2;
  while (P < n) P <<= 1;
  for (uint32_t i = t; i < P; i += blockDim.x) sk[i] = i < n ? keys[i] : ~0ull;
  __syncthreads();
  wg_bitonic_sort(sk, P);
  for (uint32_t i = t; i < n; i += blockDim.x) keys[i] = sk[i];
}
}  // namespace

void radix_sort_keys(uint64_t *keys, uint64_t n, uint32_t lo_bits, uint32_t hi_bits, SortScratch &ss,
                     hipStream_t s) {
  if (n < 2) return;
  if (n <= kSortSmallKeys) {  // one launch instead of the histogram and a pass per digit
    k_sort_small_u64<<<1, 1024, 0, s>>>(keys, static_cast<uint32_t>(n));
    TG_CHECK_LAUNCH();
    return;
  }
  radix_sort_pairs(keys, nullptr, n, lo_bits, hi_bits, ss, s);
}

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_fill_u32<<<grid_for(n, 256, 65536), 256, 0, s>>>(p, v, n);
  TG_CHECK_LAUNCH();
}
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_fill_u64<<<grid_for(n, 256, 65536), 256, 0, s>>>(p, v, n);
  TG_CHECK_LAUNCH();
}
void fill_f64(double *p, double v, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_fill_f64<<<grid_for(n, 256, 65536), 256, 0, s>>>(p, v, n);
  TG_CHECK_LAUNCH();
}

namespace {
__global__ void __launch_bounds__(256) k_copy16(uint4 *dst, const uint4 *src, uint64_t n16) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n16;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}
__global__ void k_copy1(uint8_t *dst, const uint8_t *src, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

void copy_bytes(void *dst, const void *src, uint64_t bytes, hipStream_t s) {
  if (!bytes) return;
  auto d = reinterpret_cast<uintptr_t>(dst), r = reinterpret_cast<uintptr_t>(src);
  uint64_t n16 = ((d | r) & 15) == 0 ? bytes / 16 : 0;
  if (n16) {
    k_copy16<<<grid_for(n16, 256, 1u << 16), 256, 0, s>>>(static_cast<uint4 *>(dst), static_cast<const uint4 *>(src),
                                                          n16);
    TG_CHECK_LAUNCH();
  }
  uint64_t tail = bytes - n16 * 16;
  if (tail) {
    k_copy1<<<grid_for(tail, 256), 256, 0, s>>>(static_cast<uint8_t *>(dst) + n16 * 16,
                                                 static_cast<const uint8_t *>(src) + n16 * 16, tail);
    TG_CHECK_LAUNCH();
  }
}

uint32_t LookBack::begin(uint64_t words, hipStream_t s) {
  if (misc.ensure(2, s)) fill_u32(misc.get(), 0, 2, s);
  bool fresh = state.ensure(words, s);
  if (fresh) fill_u64(state.get(), 0, state.cap, s);
  if (++epoch >= (1u << 30)) {  // wrap: old words could match again
    epoch = 1;
    fill_u64(state.get(), 0, state.cap, s);
  }
  return epoch;
}

uint32_t check_lookback(ScanScratch &ss, hipStream_t s) {
  if (!ss.lb.misc.get()) return 0;
  uint32_t e = 0;
  TG_HIP(hipMemcpyAsync(&e, ss.lb.err_word(), 4, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  if (e) fill_u32(ss.lb.err_word(), 0, 1, s);
  return e;
}

}  // namespace tg
