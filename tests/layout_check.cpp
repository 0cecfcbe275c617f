// Occ over every BWT row from the device layouts (the 32-byte wide lines and
// the 16-byte narrow lines with their superblock counts), through the same
// lookup code the kernels use (compiled for the host), against a running
// count of the BWT symbols; and the lines that the device builds from the
// BWA .bwt body (occ_line_from_bwa, the body of k_occ_lines) against the host
// construction, word for word.  Prints "ok <rows> <superblocks>" or the
// first mismatch.  Built and run by tests/test_layout.py.
#include "g3sa_amd/csrc/g3a_dev.h"
#include "g3sa_amd/csrc/g3a_index.h"
#include "g3sa_amd/csrc/g3a_layout.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace g3a;

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  HostIndex h;
  std::string err;
  if (!load_index(argv[1], h, err)) {
    std::printf("load failed: %s\n", err.c_str());
    return 1;
  }
  unpack_bwt(h);
  DevIndex x;
  std::memset(&x, 0, sizeof x);
  std::vector<uint32_t> wide, narrow, sb;
  if (!build_occ_lines(h, wide, x.super) || !build_occ_narrow(h, narrow, sb)) {
    std::printf("layout failed\n");
    return 1;
  }
  {  // the device builder's lines, word for word
    std::vector<uint32_t> dw(wide.size(), 0xdeadbeef), dn(narrow.size(), 0xdeadbeef), ds(sb.size(), 0xdeadbeef);
    const uint64_t rows = h.seq_len + 1, n_lines = (rows >> 6) + 1 + (rows >> 5) + 1;
    for (uint64_t L = 0; L < n_lines; ++L)
      occ_line_from_bwa(h.bwt_words.data(), h.seq_len, h.primary, L, dw.data(), dn.data(), ds.data());
    if (dw != wide || dn != narrow || ds != sb) {
      std::printf("device-built lines differ (wide %d narrow %d superblocks %d)\n", dw != wide, dn != narrow,
                  ds != sb);
      return 1;
    }
  }
  x.occ = wide.data();
  x.occn = narrow.data();
  x.sbc = sb.data();
  x.primary = (int64_t)h.primary;
  x.n = (int64_t)h.seq_len;
  build_counts(h, x.count);
  if (h.sa_width == 4) x.sa32 = (const uint32_t *)h.sa.data();
  else x.sa64 = (const uint64_t *)h.sa.data();
  x.sa_shift = log2_exact(h.sa_intv);
  uint64_t cnt[4] = {0, 0, 0, 0};
  const uint64_t rows = h.seq_len + 1;
  for (uint64_t p = 0; p < rows; ++p) {
    const uint8_t ch = h.bwt[(size_t)p];
    if (ch < 4) ++cnt[ch];
    int64_t ow[4];
    uint32_t on[4];
    occ4(x, (int64_t)p, ow);
    OccLineN L = occn_line(x, (uint32_t)(p >> 5));
    occ_at(x, L, (uint32_t)p, on);
    for (int c = 0; c < 4; ++c)
      if ((uint64_t)ow[c] != cnt[c] || on[c] != cnt[c]) {
        std::printf("mismatch row %llu base %d: true %llu wide %lld narrow %u\n", (unsigned long long)p, c,
                    (unsigned long long)cnt[c], (long long)ow[c], on[c]);
        return 1;
      }
    if (p % 7 == 0) {  // SA of the row by LF walks on both layouts agree
      const int64_t sw = sa_value(x, (int64_t)p, nullptr), sn = sa_value_n(x, (uint32_t)p);
      if (sw != sn) {
        std::printf("sa mismatch row %llu: wide %lld narrow %lld\n", (unsigned long long)p, (long long)sw, (long long)sn);
        return 1;
      }
    }
    const int bc = bwt_char(x, (int64_t)p);
    if (bc != (p == h.primary ? 4 : ch)) {
      std::printf("bwt_char mismatch row %llu\n", (unsigned long long)p);
      return 1;
    }
  }
  std::printf("ok %llu %zu\n", (unsigned long long)rows, sb.size() / 4);
  return 0;
}
