// AdjacencyGraph input (Ligra's text format, read by the reference's
// get_edges_from_file_adj_sym, include/io_util.h:112-145): the word
// "AdjacencyGraph", then n, m, n offsets and m neighbour ids, separated by
// whitespace.  Vertex i owns neighbours [offsets[i], offsets[i+1]) (the last
// vertex ends at m).  Edges are taken as written (the file is expected to be
// symmetric already).  Host code: the file is mapped and parsed in one pass.
#include "io.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <string>

namespace tg {

namespace {

struct Mapped {
  const char *p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap(const_cast<char *>(p), n);
    if (fd >= 0) close(fd);
  }
};

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

struct Cursor {
  const char *p, *e;
  bool word(const char *&b, size_t &len) {
    while (p < e && is_space(*p)) p++;
    if (p >= e) return false;
    b = p;
    while (p < e && !is_space(*p)) p++;
    len = static_cast<size_t>(p - b);
    return true;
  }
  // Unsigned decimal; false at end of input or on a malformed number.
  bool number(uint64_t &v) {
    const char *b;
    size_t len;
    if (!word(b, len)) return false;
    v = 0;
    for (size_t i = 0; i < len; i++) {
      if (b[i] < '0' || b[i] > '9') return false;
      v = v * 10 + static_cast<uint64_t>(b[i] - '0');
    }
    return true;
  }
};

}  // namespace

AdjFile read_adjacency_graph(const char *path) {
  Mapped f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) throw Error(-1, std::string("cannot open ") + path);
  struct stat st;
  if (fstat(f.fd, &st) != 0) throw Error(-1, std::string("cannot stat ") + path);
  f.n = static_cast<size_t>(st.st_size);
  if (f.n == 0) throw Error(-1, "Bad input file: missing header: AdjacencyGraph");
  void *m = mmap(nullptr, f.n, PROT_READ, MAP_PRIVATE, f.fd, 0);
  if (m == MAP_FAILED) throw Error(-1, std::string("cannot map ") + path);
  f.p = static_cast<const char *>(m);
  Cursor c{f.p, f.p + f.n};
  const char *b;
  size_t len;
  if (!c.word(b, len) || len != 14 || std::memcmp(b, "AdjacencyGraph", 14) != 0)
    throw Error(-1, "Bad input file: missing header: AdjacencyGraph");
  uint64_t n, mm;
  if (!c.number(n) || !c.number(mm)) throw Error(-1, "Bad input file: missing n or m");
  if (n == 0 || n > 0xFFFFFFFFull) throw Error(-1, "Bad input file: n must be in [1, 2^32)");
  AdjFile out;
  out.n = static_cast<uint32_t>(n);
  out.m = mm;
  std::vector<uint64_t> off(n);
  for (uint64_t i = 0; i < n; i++)
    if (!c.number(off[i]))
      throw Error(-1, "Bad input file: length = " + std::to_string(2 + i) + " n+m+2 = " + std::to_string(n + mm + 2));
  out.src.resize(mm);
  out.dst.resize(mm);
  for (uint64_t i = 0; i < n; i++) {
    uint64_t o = off[i], e = (i + 1 == n) ? mm : off[i + 1];
    if (o > e || e > mm) throw Error(-1, "Bad input file: offsets not increasing or beyond m");
    for (uint64_t j = o; j < e; j++) out.src[j] = static_cast<uint32_t>(i);
  }
  if (n && off[0] != 0) throw Error(-1, "Bad input file: offsets[0] must be 0");
  for (uint64_t j = 0; j < mm; j++) {
    uint64_t v;
    if (!c.number(v))
      throw Error(-1, "Bad input file: length = " + std::to_string(2 + n + j) + " n+m+2 = " + std::to_string(n + mm + 2));
    if (v >= n) throw Error(-1, "Bad input file: neighbour id out of range");
    out.dst[j] = static_cast<uint32_t>(v);
  }
  uint64_t extra;
  if (c.number(extra) || c.word(b, len))
    throw Error(-1, "Bad input file: more than n+m+2 entries (n+m+2 = " + std::to_string(n + mm + 2) + ")");
  return out;
}

}  // namespace tg
