// Device rMAT edge generator: the closed-form hash-driven generator of
// include/rmat_util.h:29-115, evaluated one edge per thread straight into
// HBM so workloads never cross PCIe.  nominal=0 divides hash32 by
// INT32_MAX exactly as rmat_util.h:95-96; nominal=1 divides by 2^32.
#include "gen.hpp"

namespace tg {

namespace {

__device__ inline uint32_t hash32(uint32_t a) {  // rmat_util.h:44-52
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

struct RmatParams {
  uint32_t n, levels, h;
  double a, ab, abc, denom;
};

__device__ inline void rmat_edge(const RmatParams &P, uint32_t i, uint32_t &x, uint32_t &y) {
  uint32_t rs = hash32((2u * i) * P.h);
  uint32_t st = hash32((2u * i + 1u) * P.h);
  x = 0;
  y = 0;
  for (uint32_t k = 0; k < P.levels; k++) {
    uint32_t half = (P.n >> k) >> 1;
    double r = static_cast<double>(hash32(rs + k * st)) / P.denom;
    if (r < P.a) {
    } else if (r < P.ab) {
      y += half;
    } else if (r < P.abc) {
      x += half;
    } else {
      x += half;
      y += half;
    }
  }
}

// flags: 1 = also write the reverse edge, 2 = drop self-loops.
__global__ void k_rmat(RmatParams P, uint64_t start, uint64_t count, int flags, uint32_t *src, uint32_t *dst,
                       unsigned long long *written) {
  for (uint64_t e0 = blockIdx.x * static_cast<uint64_t>(blockDim.x); e0 < count;
       e0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    uint64_t e = e0 + threadIdx.x;
    bool ok = e < count;
    uint32_t x = 0, y = 0;
    if (ok) rmat_edge(P, static_cast<uint32_t>(start + e), x, y);
    bool keep = ok && !((flags & 2) && x == y);
    uint32_t per = (flags & 1) ? 2u : 1u;
    uint64_t m = __ballot(keep);
    uint32_t lane = __lane_id();
    int leader = m ? __ffsll(static_cast<unsigned long long>(m)) - 1 : 0;
    unsigned long long base = 0;
    if (m && static_cast<int>(lane) == leader) base = atomicAdd(written, static_cast<unsigned long long>(__popcll(m)) * per);
    base = __shfl(base, leader, 64);
    if (keep) {
      uint64_t o = base + static_cast<uint64_t>(__popcll(m & lanemask_lt())) * per;
      src[o] = x;
      dst[o] = y;
      if (flags & 1) {
        src[o + 1] = y;
        dst[o + 1] = x;
      }
    }
  }
}

}  // namespace

uint64_t rmat_device(uint32_t n, uint32_t seed, double a, double b, double c, int nominal, uint64_t start,
                     uint64_t count, int flags, uint32_t *src, uint32_t *dst, hipStream_t s) {
  double ab = a + b, abc = a + b + c;
  if (abc > 1.0) throw Error(-1, "rMat: a + b + c > 1");
  if (n == 0 || (n & (n - 1)) != 0) throw Error(-1, "rMat: n not a power of 2");
  RmatParams P;
  P.n = n;
  P.levels = 0;
  while ((1u << P.levels) < n) P.levels++;
  // h = hash32(seed) (rmat_util.h:90)
  uint32_t h = seed;
  h = (h + 0x7ed55d16u) + (h << 12);
  h = (h ^ 0xc761c23cu) ^ (h >> 19);
  h = (h + 0x165667b1u) + (h << 5);
  h = (h + 0xd3a2646cu) ^ (h << 9);
  h = (h + 0xfd7046c5u) + (h << 3);
  h = (h ^ 0xb55a4f09u) ^ (h >> 16);
  P.h = h;
  P.a = a;
  P.ab = ab;
  P.abc = abc;
  P.denom = nominal ? 4294967296.0 : 2147483647.0;
  DevBuf<uint64_t> w;
  w.ensure(1, s);
  fill_u64(w.get(), 0, 1, s);
  if (count) {
    k_rmat<<<grid_for(count, 256), 256, 0, s>>>(P, start, count, flags, src, dst,
                                                reinterpret_cast<unsigned long long *>(w.get()));
    TG_CHECK_LAUNCH();
  }
  return read_scalar(w.get(), s);
}

}  // namespace tg
