// Host construction of the device-resident index layout from a loaded index.
#pragma once
#include "g3a_index.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace g3a {

// One 64-byte line per 64 BWT rows: 4 cumulative counts then 4 one-hot masks
// (bit 63-y marks row 64*line+y).  Rows 0..n; the '$' row sets no bit.
inline void build_occ_lines(const HostIndex &h, std::vector<uint64_t> &lines) {
  const uint64_t rows = h.seq_len + 1;
  const uint64_t n_lines = (rows >> 6) + 1;
  lines.assign((size_t)(n_lines * 8), 0);
  uint64_t c[4] = {0, 0, 0, 0};
  for (uint64_t L = 0; L < n_lines; ++L) {
    uint64_t *p = &lines[(size_t)(L * 8)];
    std::memcpy(p, c, sizeof c);
    const uint64_t b0 = L * 64;
    for (uint64_t y = 0; y < 64; ++y) {
      uint64_t r = b0 + y;
      if (r >= rows) break;
      uint8_t ch = h.bwt[(size_t)r];
      if (ch < 4) {
        p[4 + ch] |= 1ULL << (63 - y);
        ++c[ch];
      }
    }
  }
}

inline void build_counts(const HostIndex &h, int64_t count[5]) {
  for (int i = 0; i < 5; ++i) count[i] = (int64_t)h.L2[i] + 1;
}

inline int log2_exact(uint32_t v) {
  int s = 0;
  while ((1u << s) < v) ++s;
  return s;
}

}  // namespace g3a
