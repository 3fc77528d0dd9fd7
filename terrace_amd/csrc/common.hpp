// Shared definitions for the gfx950 Terrace container and traversal kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace tg {

// ---- container geometry (include/graph.h:34-39, include/btree.h:23-25) ----
constexpr uint32_t kInline = 13;          // inline neighbours per vertex block
constexpr uint32_t kMediumDegree = 1024;  // PMA tier holds degree <= 1024
constexpr uint32_t kLeafSlots = 64;       // PMA leaf = one wave of slots
constexpr uint32_t kLeafShift = 6;
constexpr uint32_t kBlk = 256;            // B-tree tier leaf block capacity
constexpr uint32_t kBlkFill = 192;        // fill target when a block splits
constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // empty PMA slot (dst)

// aux word of a vertex block: bit 63 set -> B-tree id in the low bits;
// otherwise the slot of the vertex's first PMA element, or kPmaNone.
constexpr uint64_t kAuxTree = 1ull << 63;
constexpr uint64_t kPmaNone = 0x7FFFFFFFFFFFFFFFull;

// 64-byte vertex block, same shape as the reference's vertex_block
// (include/graph.h:216-224): degree, 13 smallest neighbours, aux.
struct alignas(64) VBlock {
  uint32_t degree;
  uint32_t nbr[kInline];
  uint64_t aux;
};
static_assert(sizeof(VBlock) == 64, "vertex block must be 64 bytes");
static_assert(offsetof(VBlock, nbr) == 4 && offsetof(VBlock, aux) == 56,
              "device code reads the block as 16 words: degree, 13 neighbours, aux");

__host__ __device__ inline uint64_t make_key(uint32_t s, uint32_t d) {
  return (static_cast<uint64_t>(s) << 32) | d;
}
__host__ __device__ inline uint32_t key_src(uint64_t k) { return static_cast<uint32_t>(k >> 32); }
__host__ __device__ inline uint32_t key_dst(uint64_t k) { return static_cast<uint32_t>(k); }

// ---- error handling ------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define TG_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::tg::Error(_e == hipErrorOutOfMemory ? -2 : -3,                                \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                 \
  } while (0)

#define TG_CHECK_LAUNCH() TG_HIP(hipGetLastError())

inline unsigned grid_for(uint64_t n, unsigned block, uint64_t cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// ---- wave64 helpers ------------------------------------------------------
__device__ inline uint64_t lanemask_lt() {
  uint32_t l = __lane_id();
  return l == 0 ? 0ull : ((~0ull) >> (64 - l));
}

}  // namespace tg
