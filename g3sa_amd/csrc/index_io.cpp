// Index loading: BWA .ann/.amb/.pac/.bwt/.sa plus the optional dense .g3sa.
// The file formats follow the reference's loaders (bwa/src/bwt.c:421-462 for
// .bwt/.sa; bwa/src/fastmap.cpp:147-216 for .ann/.amb/.pac).
#include "g3a_index.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace g3a {
namespace {

bool read_file(const std::string &fn, std::vector<uint8_t> &out, std::string &err) {
  FILE *fp = std::fopen(fn.c_str(), "rb");
  if (!fp) { err = "cannot open " + fn; return false; }
  std::fseek(fp, 0, SEEK_END);
  long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  out.resize((size_t)(sz > 0 ? sz : 0));
  size_t got = out.empty() ? 0 : std::fread(out.data(), 1, out.size(), fp);
  std::fclose(fp);
  if (got != out.size()) { err = "short read of " + fn; return false; }
  return true;
}

bool load_ann(const std::string &prefix, HostIndex &idx, std::string &err) {
  std::string fn = prefix + ".ann";
  FILE *fp = std::fopen(fn.c_str(), "r");
  if (!fp) { err = "cannot open " + fn; return false; }
  long long l_pac = 0;
  int n_seqs = 0;
  unsigned seed = 0;
  if (std::fscanf(fp, "%lld %d %u", &l_pac, &n_seqs, &seed) != 3 || n_seqs < 1) {
    std::fclose(fp);
    err = "malformed " + fn;
    return false;
  }
  idx.l_pac = l_pac;
  idx.anns.resize((size_t)n_seqs);
  char line[65536];
  std::fgets(line, sizeof line, fp);  // rest of header line
  for (int i = 0; i < n_seqs; ++i) {
    Contig &c = idx.anns[(size_t)i];
    if (!std::fgets(line, sizeof line, fp)) { std::fclose(fp); err = "truncated " + fn; return false; }
    // "gi name[ anno]"
    char *p = line;
    while (*p && *p != ' ') ++p;  // skip gi
    if (*p == ' ') ++p;
    while (*p && *p != ' ' && *p != '\n' && *p != '\r') c.name.push_back(*p++);
    if (*p == ' ') {
      ++p;
      while (*p && *p != '\n' && *p != '\r') c.anno.push_back(*p++);
    }
    if (c.anno == "(null)") c.anno.clear();
    long long off = 0, len = 0;
    int namb = 0;
    if (!std::fgets(line, sizeof line, fp) || std::sscanf(line, "%lld %lld %d", &off, &len, &namb) != 3) {
      std::fclose(fp);
      err = "malformed contig record in " + fn;
      return false;
    }
    c.offset = off; c.len = len; c.n_ambs = namb;
  }
  std::fclose(fp);
  return true;
}

}  // namespace

bool load_index(const char *prefix_c, HostIndex &idx, std::string &err) {
  const std::string prefix(prefix_c);
  if (!load_ann(prefix, idx, err)) return false;
  if (!read_file(prefix + ".pac", idx.pac, err)) return false;
  if ((int64_t)idx.pac.size() < idx.l_pac / 4 + 1) { err = "pac shorter than l_pac"; return false; }
  // .bwt: header, then the body read straight into idx.bwt_words
  {
    const std::string fn = prefix + ".bwt";
    FILE *fp = std::fopen(fn.c_str(), "rb");
    if (!fp) { err = "cannot open " + fn; return false; }
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    uint64_t hdr[5];
    if (sz < 40 || std::fread(hdr, 8, 5, fp) != 5) { std::fclose(fp); err = "bwt file too short"; return false; }
    idx.primary = hdr[0];
    idx.L2[0] = 0;
    for (int i = 0; i < 4; ++i) idx.L2[i + 1] = hdr[i + 1];
    idx.seq_len = idx.L2[4];
    const uint64_t n = idx.seq_len;
    const uint64_t n_words = (uint64_t)(sz - 40) / 4;
    const uint64_t n_occ = (n + 127) / 128 + 1;
    if (n != (uint64_t)(2 * idx.l_pac) || n_words != (n + 15) / 16 + n_occ * 8) {
      std::fclose(fp);
      err = n != (uint64_t)(2 * idx.l_pac) ? "bwt length is not 2*l_pac" : "bwt size inconsistent with its header";
      return false;
    }
    idx.bwt_words.resize((size_t)n_words);
    const size_t got = std::fread(idx.bwt_words.data(), 4, (size_t)n_words, fp);
    std::fclose(fp);
    if (got != n_words) { err = "short read of " + fn; return false; }
    idx.bwt.clear();
  }
  const uint64_t n = idx.seq_len;
  // Prefer the dense sample; fall back to the BWA 32-row sample.
  std::vector<uint8_t> g3;
  std::string e2;
  if (read_file(prefix + ".g3sa", g3, e2) && g3.size() >= sizeof(G3SaHeader)) {
    G3SaHeader h;
    std::memcpy(&h, g3.data(), sizeof h);
    if (std::memcmp(h.magic, "G3SA0001", 8) != 0 || h.rows != n + 1 || (h.width != 4 && h.width != 8) ||
        g3.size() != sizeof h + h.count * h.width) {
      err = "malformed " + prefix + ".g3sa";
      return false;
    }
    idx.sa_intv = h.intv;
    idx.sa_width = h.width;
    idx.sa.assign(g3.begin() + sizeof h, g3.end());
  } else {
    std::vector<uint8_t> sa;
    if (!read_file(prefix + ".sa", sa, err)) return false;
    if (sa.size() < 56) { err = "sa file too short"; return false; }
    uint64_t sh[7];
    std::memcpy(sh, sa.data(), 56);
    const uint64_t intv = sh[5];
    if (sh[0] != idx.primary || sh[6] != n || intv == 0 || (intv & (intv - 1))) {
      err = "sa file inconsistent with bwt";
      return false;
    }
    const uint64_t n_sa = (n + intv) / intv;
    if (sa.size() != 56 + (n_sa - 1) * 8) { err = "sa file size mismatch"; return false; }
    idx.sa_intv = (uint32_t)intv;
    idx.sa_width = 8;
    idx.sa.assign((size_t)(n_sa * 8), 0);
    uint64_t row0 = n;  // SA[0] is the '$' suffix
    std::memcpy(idx.sa.data(), &row0, 8);
    std::memcpy(idx.sa.data() + 8, sa.data() + 56, (size_t)((n_sa - 1) * 8));
  }
  return true;
}

void unpack_bwt(HostIndex &idx) {
  const uint64_t n = idx.seq_len;
  const uint32_t *w = idx.bwt_words.data();
  idx.bwt.assign((size_t)(n + 1), 0);
  for (uint64_t k = 0; k < n; ++k) {
    // string position k: 128-base block k>>7 holds 8 words of counts then 8 words of bases
    uint64_t word = (k >> 7) * 16 + 8 + ((k & 127) >> 4);
    uint32_t b = (w[word] >> ((~k & 15) << 1)) & 3;
    uint64_t row = k < idx.primary ? k : k + 1;
    idx.bwt[(size_t)row] = (uint8_t)b;
  }
  idx.bwt[(size_t)idx.primary] = 4;
}

}  // namespace g3a
