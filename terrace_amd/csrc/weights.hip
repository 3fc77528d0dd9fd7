// Device hash table of edge weights (see weights.hpp).
#include "weights.hpp"

namespace tg {

namespace {

#define WS_STRIDE(i, n)                                                                              \
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < (n); \
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)

__device__ inline void ws_put(uint64_t *keys, uint32_t *vals, uint64_t mask, uint64_t key, uint32_t w) {
  uint64_t i = ws_hash(key) & mask;
  for (uint64_t probe = 0; probe <= mask; probe++) {
    unsigned long long *slot = reinterpret_cast<unsigned long long *>(keys + i);
    unsigned long long old = atomicCAS(slot, static_cast<unsigned long long>(kWsEmpty),
                                       static_cast<unsigned long long>(key));
    if (old == kWsTomb)
      old = atomicCAS(slot, static_cast<unsigned long long>(kWsTomb), static_cast<unsigned long long>(key));
    if (old == kWsEmpty || old == kWsTomb || old == key) {
      vals[i] = w;
      return;
    }
    i = (i + 1) & mask;
  }
}

__global__ void k_ws_insert(uint64_t *keys, uint32_t *vals, uint64_t mask, const uint64_t *in, const uint32_t *w,
                            const uint32_t *put, uint64_t n) {
  WS_STRIDE(i, n) if (put[i]) ws_put(keys, vals, mask, in[i], w[i]);
}

__global__ void k_ws_erase(uint64_t *keys, uint64_t mask, const uint64_t *in, uint64_t n, unsigned long long *erased) {
  WS_STRIDE(j, n) {
    const uint64_t key = in[j];
    uint64_t i = ws_hash(key) & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {
      uint64_t k = keys[i];
      if (k == key) {
        keys[i] = kWsTomb;
        atomicAdd(erased, 1ull);
        break;
      }
      if (k == kWsEmpty) break;
      i = (i + 1) & mask;
    }
  }
}

__global__ void k_ws_rehash(const uint64_t *okeys, const uint32_t *ovals, uint64_t ocap, uint64_t *keys,
                            uint32_t *vals, uint64_t mask) {
  WS_STRIDE(i, ocap) {
    uint64_t k = okeys[i];
    if (k != kWsEmpty && k != kWsTomb) ws_put(keys, vals, mask, k, ovals[i]);
  }
}

__global__ void k_ws_lookup(WView v, const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t *out) {
  WS_STRIDE(i, n) out[i] = v.lookup(src[i], dst[i]);
}

}  // namespace

void WeightStore::reserve(uint64_t need, hipStream_t s) {
  // keys + tombstones stay at or below 70% of the slots
  if (cap_ && (need + tomb_) * 10 <= cap_ * 7) return;
  uint64_t cap = 1024;
  while (cap * 7 < need * 10 * 2) cap <<= 1;  // about 35% full after the rehash
  DevBuf<uint64_t> nk;
  DevBuf<uint32_t> nv;
  nk.ensure(cap, s);
  nv.ensure(cap, s);
  fill_u64(nk.get(), kWsEmpty, cap, s);
  if (cap_) {
    k_ws_rehash<<<grid_for(cap_, 256), 256, 0, s>>>(keys_.get(), vals_.get(), cap_, nk.get(), nv.get(), cap - 1);
    TG_CHECK_LAUNCH();
  }
  TG_HIP(hipStreamSynchronize(s));
  std::swap(keys_.ptr, nk.ptr);
  std::swap(keys_.cap, nk.cap);
  std::swap(vals_.ptr, nv.ptr);
  std::swap(vals_.cap, nv.cap);
  cap_ = cap;
  tomb_ = 0;
}

void WeightStore::insert(const uint64_t *keys, const uint32_t *w, const uint32_t *put, uint64_t n, uint64_t new_count,
                         hipStream_t s) {
  if (n == 0 || new_count == 0) return;
  reserve(live_ + new_count, s);
  k_ws_insert<<<grid_for(n, 256), 256, 0, s>>>(keys_.get(), vals_.get(), cap_ - 1, keys, w, put, n);
  TG_CHECK_LAUNCH();
  live_ += new_count;
}

void WeightStore::erase(const uint64_t *keys, uint64_t n, hipStream_t s) {
  if (n == 0 || cap_ == 0) return;
  cnt_.ensure(1, s);
  fill_u64(reinterpret_cast<uint64_t *>(cnt_.get()), 0, 1, s);
  k_ws_erase<<<grid_for(n, 256), 256, 0, s>>>(keys_.get(), cap_ - 1, keys, n, cnt_.get());
  TG_CHECK_LAUNCH();
  uint64_t e = read_scalar(reinterpret_cast<const uint64_t *>(cnt_.get()), s);
  live_ -= e;
  tomb_ += e;
}

void WeightStore::lookup(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t *out, hipStream_t s) const {
  if (n == 0) return;
  if (cap_ == 0) {
    fill_u32(out, kWsMissing, n, s);
    return;
  }
  k_ws_lookup<<<grid_for(n, 256), 256, 0, s>>>(view(), src, dst, n, out);
  TG_CHECK_LAUNCH();
}

}  // namespace tg
