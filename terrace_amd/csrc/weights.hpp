// Edge weights of a weighted graph (the reference's WEIGHTED build,
// include/graph.h:34-38, 76-86): one uint32 per directed edge, held in a
// device hash table keyed by (src << 32 | dst) beside the degree-tiered
// container.  The container's update and traversal kernels stay unchanged;
// the weighted Bellman-Ford op looks a weight up per relaxed edge.
//
// Open addressing with linear probing; slots hold kEmpty, kTomb (a deleted
// key) or a key.  Inserts claim a slot with atomicCAS, so one launch may
// insert many distinct keys.  The table grows (rehash) when keys plus
// tombstones pass 70% of the slots.
#pragma once

#include <cstdint>

#include "prims.hpp"

namespace tg {

constexpr uint64_t kWsEmpty = ~0ull;
constexpr uint64_t kWsTomb = ~0ull - 1;
constexpr uint32_t kWsMissing = 0xFFFFFFFFu;  // lookup result for an absent edge

// MurmurHash3's 64-bit finaliser (fmix64, public domain) spreads the key.
__host__ __device__ inline uint64_t ws_hash(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

struct WView {
  const uint64_t *keys = nullptr;
  const uint32_t *vals = nullptr;
  uint64_t mask = 0;
  __device__ uint32_t lookup(uint32_t s, uint32_t d) const {
    const uint64_t key = (static_cast<uint64_t>(s) << 32) | d;
    uint64_t i = ws_hash(key) & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {
      uint64_t k = keys[i];
      if (k == key) return vals[i];
      if (k == kWsEmpty) return kWsMissing;
      i = (i + 1) & mask;
    }
    return kWsMissing;
  }
};

class WeightStore {
 public:
  WeightStore() = default;
  // insert keys[i] -> w[i] where put[i] != 0 (keys absent from the table)
  void insert(const uint64_t *keys, const uint32_t *w, const uint32_t *put, uint64_t n, uint64_t new_count,
              hipStream_t s);
  // delete the keys present in the table
  void erase(const uint64_t *keys, uint64_t n, hipStream_t s);
  // out[i] = weight of (src[i], dst[i]) or kWsMissing
  void lookup(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t *out, hipStream_t s) const;
  WView view() const { return WView{keys_.get(), vals_.get(), cap_ ? cap_ - 1 : 0}; }
  uint64_t live() const { return live_; }
  uint64_t bytes() const { return cap_ * 12; }

 private:
  DevBuf<uint64_t> keys_;
  DevBuf<uint32_t> vals_;
  DevBuf<unsigned long long> cnt_;
  uint64_t cap_ = 0, live_ = 0, tomb_ = 0;
  void reserve(uint64_t need, hipStream_t s);
};

}  // namespace tg
