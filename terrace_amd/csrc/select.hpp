// Predicate-fused stream compaction: F::pred(i) is evaluated inside the
// count and the scatter kernels, so no flag array and no separate flag
// kernel are needed.  out positions follow index order (stable).
//
//   struct F { __device__ bool pred(uint64_t i) const;
//              __device__ void emit(uint64_t i, uint64_t pos) const; };
//   select(f, n, total_dev, scratch, stream)   // *total_dev = count
#pragma once

#include <type_traits>
#include <utility>

#include "prims.hpp"

namespace tg {

namespace sel {

// F::finish(total), when F has it, runs once after the last position is
// known (last block, one thread): e.g. a terminator after the output.
template <class F, class = void>
struct has_finish : std::false_type {};
template <class F>
struct has_finish<F, std::void_t<decltype(std::declval<const F &>().finish(uint64_t{}))>> : std::true_type {};

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // 2048, items blocked per thread

template <class F>
__global__ void __launch_bounds__(kBlock) k_count(F f, uint64_t n, uint64_t *sums) {
  __shared__ uint32_t lds[kBlock / 64];
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
    if (idx < n && f.pred(idx)) c++;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) sums[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// Items are visited in wave-strided order (coalesced); ranks come from
// ballots, so the emitted order is index order within the tile.
template <class F>
__global__ void __launch_bounds__(kBlock) k_emit(F f, uint64_t n, const uint64_t *offs, uint64_t *total) {
  __shared__ uint32_t wcnt[kItems][kBlock / 64];
  __shared__ uint32_t wpre[kItems][kBlock / 64];
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  uint64_t base = static_cast<uint64_t>(blockIdx.x) * kTile;
  bool p[kItems];
  uint32_t r[kItems];
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
    p[i] = idx < n && f.pred(idx);
    uint64_t m = __ballot(p[i]);
    r[i] = __popcll(m & lt);
    if ((threadIdx.x & 63) == 0) wcnt[i][w] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < kItems; i++)
      for (int ww = 0; ww < kBlock / 64; ww++) {
        wpre[i][ww] = run;
        run += wcnt[i][ww];
      }
    if (blockIdx.x == gridDim.x - 1) {
      const uint64_t t = (offs ? offs[blockIdx.x] : 0) + run;
      if (total) *total = t;
      if constexpr (has_finish<F>::value) f.finish(t);
    }
  }
  __syncthreads();
  uint64_t boff = offs ? offs[blockIdx.x] : 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    if (p[i]) {
      uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
      f.emit(idx, boff + wpre[i][w] + r[i]);
    }
  }
}

// Single pass: the tile's count is published and the block's offset found
// by decoupled look-back, then the tile is emitted (one launch for any n).
template <class F>
__global__ void __launch_bounds__(kBlock) k_select_lb(F f, uint64_t n, uint64_t *total, uint64_t *state,
                                                      uint32_t *misc, uint32_t *err, uint32_t nblocks,
                                                      uint32_t epoch) {
  __shared__ uint32_t wcnt[kItems][kBlock / 64];
  __shared__ uint32_t wpre[kItems][kBlock / 64];
  __shared__ uint32_t s_ticket;
  __shared__ uint64_t s_off;
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  uint64_t base = static_cast<uint64_t>(b) * kTile;
  bool p[kItems];
  uint32_t r[kItems];
  // all predicates first, so their memory accesses overlap
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
    p[i] = idx < n && f.pred(idx);
  }
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t m = __ballot(p[i]);
    r[i] = __popcll(m & lt);
    if ((threadIdx.x & 63) == 0) wcnt[i][w] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // wave 0: the tile's prefixes, then the look-back
    uint32_t run = 0;
    for (int i = 0; i < kItems; i++)
      for (int ww = 0; ww < kBlock / 64; ww++) {
        if (threadIdx.x == 0) wpre[i][ww] = run;
        run += wcnt[i][ww];
      }
    const uint64_t off = lbk::lookback_wave(state, 1, b, 0, err, epoch, run);
    if (threadIdx.x == 0) {
      s_off = off;
      if (b == nblocks - 1) {
        if (total) *total = off + run;
        if constexpr (has_finish<F>::value) f.finish(off + run);
      }
    }
  }
  __syncthreads();
  const uint64_t boff = s_off;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    if (p[i]) {
      uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
      f.emit(idx, boff + wpre[i][w] + r[i]);
    }
  }
}

// Two outputs in one pass: F::pick(i) gives 0 (neither), 1 or 2, and
// F::emit(i, which, pos) writes to output `which` (1 or 2).  Each output
// has its own look-back column (state stride 2); wave 0 walks column 0 and
// wave 1 column 1 at the same time.
template <class F>
__global__ void __launch_bounds__(kBlock) k_select2_lb(F f, uint64_t n, uint64_t *total1, uint64_t *total2,
                                                       uint64_t *state, uint32_t *misc, uint32_t *err,
                                                       uint32_t nblocks, uint32_t epoch) {
  __shared__ uint32_t wcnt[2][kItems][kBlock / 64];
  __shared__ uint32_t wpre[2][kItems][kBlock / 64];
  __shared__ uint32_t s_ticket, s_run[2];
  __shared__ uint64_t s_off[2];
  const uint32_t b = lbk::ticket(misc, err, nblocks, &s_ticket);
  if (b >= nblocks) return;
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  uint64_t base = static_cast<uint64_t>(b) * kTile;
  uint32_t c[kItems], r[kItems];
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
    c[i] = idx < n ? f.pick(idx) : 0u;
  }
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    const uint64_t m1 = __ballot(c[i] == 1), m2 = __ballot(c[i] == 2);
    r[i] = static_cast<uint32_t>(__popcll((c[i] == 2 ? m2 : m1) & lt));
    if ((threadIdx.x & 63) == 0) {
      wcnt[0][i][w] = static_cast<uint32_t>(__popcll(m1));
      wcnt[1][i][w] = static_cast<uint32_t>(__popcll(m2));
    }
  }
  __syncthreads();
  if (threadIdx.x < 128) {  // wave 0 walks column 0, wave 1 column 1
    const uint32_t col = threadIdx.x >> 6;
    const bool lead = (threadIdx.x & 63) == 0;
    uint32_t run = 0;
    for (int i = 0; i < kItems; i++)
      for (int ww = 0; ww < kBlock / 64; ww++) {
        if (lead) wpre[col][i][ww] = run;
        run += wcnt[col][i][ww];
      }
    const uint64_t off = lbk::lookback_wave(state, 2, b, col, err, epoch, run);
    if (lead) {
      s_off[col] = off;
      s_run[col] = run;
    }
  }
  __syncthreads();
  if (b == nblocks - 1 && threadIdx.x == 0) {
    if (total1) *total1 = s_off[0] + s_run[0];
    if (total2) *total2 = s_off[1] + s_run[1];
  }
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    if (c[i]) {
      const uint32_t col = c[i] - 1;
      uint64_t idx = base + static_cast<uint64_t>(i) * kBlock + threadIdx.x;
      f.emit(idx, c[i], s_off[col] + wpre[col][i][w] + r[i]);
    }
  }
}

}  // namespace sel

// Compacts the indices with f.pred into f.emit positions; *total_dev gets
// the count (device scalar, not synchronised).
template <class F>
void select(const F &f, uint64_t n, uint64_t *total_dev, ScanScratch &ss, hipStream_t s) {
  if (n == 0) {
    fill_u64(total_dev, 0, 1, s);
    return;
  }
  uint64_t nblocks = (n + sel::kTile - 1) / sel::kTile;
  if (nblocks == 1) {
    sel::k_emit<F><<<1, sel::kBlock, 0, s>>>(f, n, nullptr, total_dev);
    TG_CHECK_LAUNCH();
    return;
  }
  if (ss.lookback && n <= 0xFFFFFFFFull) {  // counts fit the look-back word
    uint32_t nb = static_cast<uint32_t>(nblocks);
    uint32_t epoch = ss.lb.begin(nb, s);
    sel::k_select_lb<F><<<nb, sel::kBlock, 0, s>>>(f, n, total_dev, ss.lb.state.get(), ss.lb.misc.get(),
                                                   ss.lb.err_word(), nb, epoch);
    TG_CHECK_LAUNCH();
    return;
  }
  DevBuf<uint64_t> &sums = ss.levels[5];
  sums.ensure(nblocks + 1, s);
  sel::k_count<F><<<static_cast<unsigned>(nblocks), sel::kBlock, 0, s>>>(f, n, sums.get());
  TG_CHECK_LAUNCH();
  exclusive_scan_u64(sums.get(), sums.get(), nblocks, nullptr, ss, s);
  sel::k_emit<F><<<static_cast<unsigned>(nblocks), sel::kBlock, 0, s>>>(f, n, sums.get(), total_dev);
  TG_CHECK_LAUNCH();
}

// Two compactions in one pass (see k_select2_lb); the totals go to two
// device scalars.  Needs the look-back scratch.
template <class F>
void select2(const F &f, uint64_t n, uint64_t *total1, uint64_t *total2, ScanScratch &ss, hipStream_t s) {
  if (!ss.lookback || n > 0xFFFFFFFFull) throw Error(-1, "select2: needs the look-back scratch and n < 2^32");
  if (n == 0) {
    fill_u64(total1, 0, 1, s);
    fill_u64(total2, 0, 1, s);
    return;
  }
  const uint32_t nb = static_cast<uint32_t>((n + sel::kTile - 1) / sel::kTile);
  const uint32_t epoch = ss.lb.begin(2ull * nb, s);
  sel::k_select2_lb<F><<<nb, sel::kBlock, 0, s>>>(f, n, total1, total2, ss.lb.state.get(), ss.lb.misc.get(),
                                                  ss.lb.err_word(), nb, epoch);
  TG_CHECK_LAUNCH();
}

// Several device scalars in one copy (synchronises the stream).
inline void read_scalars(const uint64_t *dev, uint64_t *host, int count, hipStream_t s) {
  TG_HIP(hipMemcpyAsync(host, dev, count * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
}

}  // namespace tg
