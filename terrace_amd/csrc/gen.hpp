#pragma once

#include "prims.hpp"

namespace tg {

// Writes rMat<uint32_t>(i) for i in [start, start+count) into device
// arrays (flags: 1 = add the reverse edge right after, 2 = drop
// self-loops).  Output order is not the index order; returns the number of
// (src,dst) entries written.
uint64_t rmat_device(uint32_t n, uint32_t seed, double a, double b, double c, int nominal, uint64_t start,
                     uint64_t count, int flags, uint32_t *src, uint32_t *dst, hipStream_t s);

}  // namespace tg
