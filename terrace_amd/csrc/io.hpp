// AdjacencyGraph text input (include/io_util.h:112-145).
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace tg {

struct AdjFile {
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> src, dst;  // edge j = (src[j], dst[j]) in file order
};

// Parses the file; throws Error(-1, ...) on a malformed file (the reference
// exits in those cases).
AdjFile read_adjacency_graph(const char *path);

}  // namespace tg
