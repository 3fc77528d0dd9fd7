// Device primitives: grow-only scratch buffers, exclusive scan, LSD radix
// sort of 64-bit keys.  All work is issued on the caller's stream.
#pragma once

#include "common.hpp"

namespace tg {

// Grow-only device buffer.  ensure() keeps contents only when keep=true.
template <typename T>
struct DevBuf {
  T *ptr = nullptr;
  uint64_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
  // Ensure room for n elements; returns true if the storage moved.
  bool ensure(uint64_t n, hipStream_t s = nullptr, bool keep = false, uint64_t keep_n = 0) {
    if (n <= cap && ptr) return false;
    uint64_t nc = n < 16 ? 16 : n + n / 4;
    T *np = nullptr;
    TG_HIP(hipMalloc(reinterpret_cast<void **>(&np), nc * sizeof(T)));
    if (keep && ptr && keep_n) {
      TG_HIP(hipMemcpyAsync(np, ptr, keep_n * sizeof(T), hipMemcpyDeviceToDevice, s));
      TG_HIP(hipStreamSynchronize(s));
    }
    if (ptr) {
      TG_HIP(hipStreamSynchronize(s));
      TG_HIP(hipFree(ptr));
    }
    ptr = np;
    cap = nc;
    return true;
  }
  T *get() const { return ptr; }
  operator T *() const { return ptr; }
};

// Ascending bitonic sort of s[0, P) in LDS by the whole workgroup: P a power
// of two, s filled (padded with ~0) and a barrier after the writes; every
// thread calls it, and it returns after a barrier.  With at most one key
// per thread a thread keeps its key in a register: the stages whose partner
// is in the same wave (distance below 64) exchange lanes, and only the
// wider ones go through s with a barrier pair (10 of the 55 stages of a
// 1024-key sort, which takes about 9 us instead of 14).  More keys than
// threads: one compare-exchange per thread and stage over s.
__device__ inline void wg_bitonic_sort(uint64_t *s, uint32_t P) {
  const uint32_t t = threadIdx.x;
  if (P <= blockDim.x) {  // uniform
    uint64_t v = t < P ? s[t] : ~0ull;
    for (uint32_t k = 2; k <= P; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        uint64_t p;
        if (j < 64) {
          p = __shfl_xor(v, static_cast<int>(j), 64);
        } else {
          __syncthreads();  // the previous stage's reads of s are done
          if (t < P) s[t] = v;
          __syncthreads();
          p = t < P ? s[t ^ j] : ~0ull;
        }
        // the key at the lower index keeps the smaller one when ascending
        const bool lower = (t & j) == 0, up = (t & k) == 0;
        v = (lower == up) ? (v < p ? v : p) : (v < p ? p : v);
      }
    }
    __syncthreads();
    if (t < P) s[t] = v;
  } else {
    for (uint32_t k = 2; k <= P; k <<= 1) {
      for (uint32_t j = k >> 1; j > 0; j >>= 1) {
        for (uint32_t i = t; i < P; i += blockDim.x) {
          const uint32_t ixj = i ^ j;
          if (ixj > i) {
            const uint64_t a = s[i], b = s[ixj];
            if ((a > b) == ((i & k) == 0)) {
              s[i] = b;
              s[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();
}

// Single-pass (decoupled look-back) state.  Blocks take tickets in the
// order they start, so a block only waits on blocks already running; each
// publishes its own total before it waits.  Words carry the launch's epoch,
// so stale words from earlier launches never match and the array needs no
// clearing between launches.  Every wait is bounded: a block that gives up
// sets misc[1] and the caller reports it (check_lookback).
struct LookBack {
  DevBuf<uint64_t> state;
  DevBuf<uint32_t> misc;  // [0] ticket counter, [1] error (a wait gave up)
  uint32_t *err = nullptr;  // error word elsewhere (null: misc + 1)
  uint32_t epoch = 0;
  uint32_t *err_word() { return err ? err : misc.get() + 1; }
  // Room for `words` state words; returns the epoch for the next launch.
  uint32_t begin(uint64_t words, hipStream_t s);
};

struct ScanScratch {
  DevBuf<uint64_t> levels[6];
  bool lookback = false;  // single-pass scan / select / sort on this scratch
  LookBack lb;
};

namespace lbk {

constexpr uint32_t kSpinLimit = 1u << 22;  // bounded wait (about 10^8 cycles)

// One 64-bit word per published value: epoch << 34 | status << 32 | value
// (status 1 = the block's own total, 2 = inclusive prefix).  The word is
// the whole message, so relaxed agent-scope atomics suffice: they read and
// write the device-coherent copy (XCDs have separate L2 caches) without
// cache maintenance.
__device__ inline uint64_t load(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t word(uint32_t epoch, uint32_t status, uint32_t value) {
  return (static_cast<uint64_t>(epoch) << 34) | (static_cast<uint64_t>(status) << 32) | value;
}

// Launches of at most this many blocks use the block index as their order:
// the dispatcher starts workgroups in index order, so a block only waits on
// blocks already started, and the ticket's atomic round trip is saved.
constexpr uint32_t kOrderedBlocks = 1024;

// Block ticket in start order (one thread takes it, the block shares it).
// The block that takes the last ticket resets the counter for the next
// launch on the stream.  Returns nblocks when the counter was inconsistent
// (the caller skips its work and the error word is set).
__device__ inline uint32_t ticket(uint32_t *misc, uint32_t *err, uint32_t nblocks, uint32_t *lds_slot) {
  if (nblocks <= kOrderedBlocks) return blockIdx.x;
  if (threadIdx.x == 0) {
    uint32_t v = atomicAdd(&misc[0], 1u);
    if (v == nblocks - 1) atomicExch(&misc[0], 0u);
    if (v >= nblocks) {
      atomicOr(err, 2u);
      v = nblocks;
    }
    *lds_slot = v;
  }
  __syncthreads();
  return *lds_slot;
}

// Publishes agg for slot (b, col) and returns the sum of the values before
// it (slots (b - 1, col), (b - 2, col), ... down to block 0).  Totals must
// fit 32 bits.  Run by one thread per slot, reading kLbWindow predecessors
// with independent loads per round trip (for launches with many slots per
// block, e.g. one per radix digit).
constexpr int kLbWindow = 4;

__device__ inline uint32_t lookback_window(uint64_t *state, uint64_t stride, uint32_t b, uint64_t col,
                                           uint32_t *err, uint32_t epoch, uint32_t agg) {
  uint64_t *me = state + static_cast<uint64_t>(b) * stride + col;
  if (b == 0) {
    store(me, word(epoch, 2, agg));
    return 0;
  }
  store(me, word(epoch, 1, agg));
  uint32_t excl = 0;
  int64_t p = static_cast<int64_t>(b) - 1;
  uint32_t spins = 0;
  const uint64_t tag = static_cast<uint64_t>(epoch) << 34;
  while (p >= 0) {
    uint64_t f[kLbWindow];
#pragma unroll
    for (int j = 0; j < kLbWindow; j++)
      f[j] = p - j >= 0 ? load(state + static_cast<uint64_t>(p - j) * stride + col) : 0;
    // consume in order up to the first word not ready or the first inclusive
    bool done = false, stalled = false;
    int used = 0;
#pragma unroll
    for (int j = 0; j < kLbWindow; j++) {
      if (done || stalled || p - j < 0) continue;
      const uint32_t st = static_cast<uint32_t>(f[j] >> 32) & 3u;
      if ((f[j] & ~((1ull << 34) - 1)) != tag || st == 0) {
        stalled = true;
        continue;
      }
      excl += static_cast<uint32_t>(f[j]);
      used++;
      if (st == 2) done = true;
    }
    if (done) break;
    p -= used;
    if (stalled) {
      if (++spins > kSpinLimit) {
        atomicOr(err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  store(me, word(epoch, 2, excl + agg));
  return excl;
}

// The same for one slot per block, run by all 64 lanes of one wave (same
// arguments in every lane; the result is the same in every lane): lane l
// reads predecessor top - l, so one round trip covers 64 predecessors; a
// ballot finds the nearest inclusive prefix and a wave sum adds the words
// up to it.
__device__ inline uint32_t lookback_wave(uint64_t *state, uint64_t stride, uint32_t b, uint64_t col,
                                         uint32_t *err, uint32_t epoch, uint32_t agg) {
  const uint32_t lane = threadIdx.x & 63;
  uint64_t *me = state + static_cast<uint64_t>(b) * stride + col;
  if (b == 0) {
    if (lane == 0) store(me, word(epoch, 2, agg));
    return 0;
  }
  if (lane == 0) store(me, word(epoch, 1, agg));
  uint32_t excl = 0;
  int64_t top = static_cast<int64_t>(b) - 1;
  uint32_t spins = 0;
  const uint64_t tag = static_cast<uint64_t>(epoch) << 34;
  while (true) {
    const int64_t p = top - static_cast<int64_t>(lane);
    bool ready = true;
    uint32_t st = 2, val = 0;  // past block 0: never consumed (block 0 is inclusive)
    if (p >= 0) {
      const uint64_t f = load(state + static_cast<uint64_t>(p) * stride + col);
      st = static_cast<uint32_t>(f >> 32) & 3u;
      ready = (f & ~((1ull << 34) - 1)) == tag && st != 0;
      val = static_cast<uint32_t>(f);
    }
    const uint64_t inc = __ballot(ready && st == 2);
    const uint64_t stalled = __ballot(!ready);
    const int first = inc ? __ffsll(static_cast<unsigned long long>(inc)) - 1 : 63;
    const uint64_t need = first >= 63 ? ~0ull : ((1ull << (first + 1)) - 1);  // lanes 0..first
    if (stalled & need) {
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, 1u);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    uint32_t v = static_cast<int>(lane) <= first ? val : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    excl += v;
    if (inc) break;
    top -= 64;
  }
  if (lane == 0) store(me, word(epoch, 2, excl + agg));
  return excl;
}

}  // namespace lbk

// Nonzero when a single-pass kernel on this scratch gave up a wait
// (synchronises the stream); clears the flag.
uint32_t check_lookback(ScanScratch &ss, hipStream_t s);

// out[i] = sum(in[0..i)), returns nothing; *total_dev (optional) gets the sum,
// and so does *total64 (optional, as a 64-bit word).  in and out may alias.
// n may be 0.
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total_dev,
                        ScanScratch &ss, hipStream_t s, uint64_t *total64 = nullptr);
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                        ScanScratch &ss, hipStream_t s);

struct SortScratch {
  DevBuf<uint64_t> alt;
  DevBuf<uint32_t> alt_vals;
  DevBuf<uint32_t> hist;
  ScanScratch scan;
};

// Stable LSD radix sort of keys in place over the bit ranges that vary:
// dst bits [0, lo_bits) and src bits [32, 32 + hi_bits) (every other bit
// zero).  8-bit digits.  At most 4096 keys are sorted by one workgroup in
// LDS instead (the same order: keys carry no values).
void radix_sort_keys(uint64_t *keys, uint64_t n, uint32_t lo_bits, uint32_t hi_bits,
                     SortScratch &ss, hipStream_t s);

// Same, carrying one u32 value per key (stable).
void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t n, uint32_t lo_bits, uint32_t hi_bits,
                      SortScratch &ss, hipStream_t s);

// Fill helpers (kernels, not hipMemset, so they are safe under capture).
void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
void fill_f64(double *p, double v, uint64_t n, hipStream_t s);

// Device-to-device copy with a plain 16-byte-per-lane kernel (its HBM
// traffic is exactly bytes read + bytes written).
void copy_bytes(void *dst, const void *src, uint64_t bytes, hipStream_t s);

// Read one device scalar (synchronises the stream).
template <typename T>
T read_scalar(const T *dev, hipStream_t s) {
  T h{};
  TG_HIP(hipMemcpyAsync(&h, dev, sizeof(T), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace tg
