// Host construction of the device-resident index layout from a loaded index.
#pragma once
#include "g3a_dev.h"
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

// Narrow layout, for texts of fewer than 2^32 rows: one 16-byte line per 32
// BWT rows, so that an Occ lookup is a single 16-byte load.  Bytes 0-7 (a
// little-endian uint64): the counts of bases 0, 1 and 2 in the rows before
// the line, relative to the line's 2^21-row superblock, at bits 0-20, 21-41
// and 42-62; base 3's count follows from the row count.  Bytes 8-15: the 32
// rows' symbols at 2 bits each as one uint64 (row y at bits 63-2y and 62-2y),
// the '$' row stored as 0 and never counted.  sb holds four uint32 per
// superblock: each base's count in the rows before the superblock.
inline bool build_occ_narrow(const HostIndex &h, std::vector<uint32_t> &lines, std::vector<uint32_t> &sb) {
  const uint64_t rows = h.seq_len + 1;
  if (rows >= (1ULL << 32)) return false;
  const uint64_t n_lines = (rows >> 5) + 1, n_sb = (rows >> kNarrowSbShift) + 1;
  lines.assign((size_t)(n_lines * 4), 0);
  sb.assign((size_t)(n_sb * 4), 0);
  uint64_t c[4] = {0, 0, 0, 0}, base[4] = {0, 0, 0, 0};
  for (uint64_t L = 0; L < n_lines; ++L) {
    const uint64_t b0 = L * 32;
    if ((b0 & ((1ULL << kNarrowSbShift) - 1)) == 0) {
      const uint64_t s = b0 >> kNarrowSbShift;
      for (int k = 0; k < 4; ++k) {
        sb[(size_t)(s * 4 + k)] = (uint32_t)c[k];
        base[k] = c[k];
      }
    }
    const uint64_t packed = (c[0] - base[0]) | (c[1] - base[1]) << 21 | (c[2] - base[2]) << 42;
    uint64_t w = 0;
    for (uint64_t y = 0; y < 32; ++y) {
      const uint64_t r = b0 + y;
      if (r >= rows) break;
      const uint8_t ch = h.bwt[(size_t)r];
      if (ch < 4) {
        w |= (uint64_t)ch << (62 - 2 * y);
        ++c[ch];
      }
    }
    std::memcpy(&lines[(size_t)(L * 4)], &packed, 8);
    std::memcpy(&lines[(size_t)(L * 4 + 2)], &w, 8);
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
