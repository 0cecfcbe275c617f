// Host construction of the device-resident index layout from a loaded index.
#pragma once
#include "g3a_index.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace g3a {

// One 32-byte line per 64 BWT rows: four uint32 counts of each base in the
// rows before the line (relative to the line's 2^32-row superblock), then the
// 64 rows' BWT symbols at 2 bits each as two uint64 words (row y of the line
// in word y/32, bits 63-2(y%32) and 62-2(y%32), the first row highest).  The
// '$' row is stored as symbol 0 and is never counted.  super[s][c] holds the
// count of base c in rows before superblock s.  Rows 0..n.
inline bool build_occ_lines(const HostIndex &h, std::vector<uint32_t> &lines, int64_t super[4][4]) {
  const uint64_t rows = h.seq_len + 1;
  const uint64_t n_lines = (rows >> 6) + 1;
  if (rows > (4ULL << 32)) return false;
  lines.assign((size_t)(n_lines * 8), 0);
  uint64_t c[4] = {0, 0, 0, 0}, base[4] = {0, 0, 0, 0};
  std::memset(super, 0, sizeof(int64_t) * 16);
  for (uint64_t L = 0; L < n_lines; ++L) {
    const uint64_t b0 = L * 64;
    if ((b0 & 0xffffffffULL) == 0) {  // superblock start
      const int sb = (int)(b0 >> 32);
      for (int k = 0; k < 4; ++k) {
        super[sb][k] = (int64_t)c[k];
        base[k] = c[k];
      }
    }
    uint32_t *p = &lines[(size_t)(L * 8)];
    for (int k = 0; k < 4; ++k) p[k] = (uint32_t)(c[k] - base[k]);
    uint64_t w[2] = {0, 0};
    for (uint64_t y = 0; y < 64; ++y) {
      uint64_t r = b0 + y;
      if (r >= rows) break;
      uint8_t ch = h.bwt[(size_t)r];
      if (ch < 4) {
        w[y >> 5] |= (uint64_t)ch << (62 - 2 * (y & 31));
        ++c[ch];
      }
    }
    std::memcpy(p + 4, w, sizeof w);
  }
  return true;
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
