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

struct ScanScratch {
  DevBuf<uint64_t> levels[6];
};

// out[i] = sum(in[0..i)), returns nothing; *total_dev (optional) gets the sum.
// in and out may alias.  n may be 0.
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *total_dev,
                        ScanScratch &ss, hipStream_t s);
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                        ScanScratch &ss, hipStream_t s);

struct SortScratch {
  DevBuf<uint64_t> alt;
  DevBuf<uint32_t> alt_vals;
  DevBuf<uint32_t> hist;
  ScanScratch scan;
};

// Stable LSD radix sort of keys in place over the bit ranges that vary:
// dst bits [0, lo_bits) and src bits [32, 32 + hi_bits).  8-bit digits.
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
