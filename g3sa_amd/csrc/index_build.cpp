// Index construction for the BWA-MEM hot path.
//
// Writes, for a FASTA reference and an output prefix:
//   <prefix>.pac .ann .amb  BWA "bntseq" files: forward strand packed 2 bits
//                           per base, contig table, ambiguous-base holes
//                           (layout as read by bwa/src/fastmap.cpp:147-216 and
//                           bwa/src/bwt.c:385-462 of the reference)
//   <prefix>.bwt            BWA FM-index of fwd+revcomp text: primary, L2[1..4],
//                           then per 128 rows 4 x uint64 Occ + 128 x 2-bit BWT
//   <prefix>.sa             BWA sampled suffix array (every 32nd row)
//   <prefix>.counts         bwa-mem2 count tables (bwa/src/FMI_search.cpp:248-259)
//   <prefix>.g3sa           dense suffix-array sample used by the GPU SA lookup
//
// The suffix array is built here by SA-IS (sais.h); `bwa index` itself has no
// source in the reference, so only its file formats are followed.
#include "g3a_index.h"
#include "sais.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace g3a {
namespace {

struct FaContig {
  std::string name, anno;
  std::vector<uint8_t> seq;  // codes 0..3, 4 = ambiguous
};

int nt4(int c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;
    default: return 4;
  }
}

bool read_fasta(const char *fn, std::vector<FaContig> &out, std::string &err) {
  FILE *fp = std::fopen(fn, "r");
  if (!fp) { err = std::string("cannot open ") + fn; return false; }
  std::vector<char> buf(1 << 20);
  while (std::fgets(buf.data(), (int)buf.size(), fp)) {
    char *p = buf.data();
    if (*p == '>') {
      out.emplace_back();
      ++p;
      while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') out.back().name.push_back(*p++);
      while (*p == ' ' || *p == '\t') ++p;
      while (*p && *p != '\n' && *p != '\r') out.back().anno.push_back(*p++);
    } else if (!out.empty()) {
      for (; *p && *p != '\n' && *p != '\r'; ++p)
        if (!std::isspace((unsigned char)*p)) out.back().seq.push_back((uint8_t)nt4(*p));
    }
  }
  std::fclose(fp);
  if (out.empty()) { err = "no sequences in FASTA"; return false; }
  return true;
}

// Deterministic generator for the bases that replace ambiguous characters.
struct HoleRng {
  uint64_t x;
  explicit HoleRng(uint64_t s) : x(s) {}
  uint32_t next() {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    return (uint32_t)x;
  }
};

bool write_all(const std::string &fn, const void *p, size_t n, std::string &err) {
  FILE *fp = std::fopen(fn.c_str(), "wb");
  if (!fp) { err = "cannot write " + fn; return false; }
  bool ok = std::fwrite(p, 1, n, fp) == n;
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) err = "short write to " + fn;
  return ok;
}

template <typename I>
bool build_sa(const std::vector<uint8_t> &text, std::vector<I> &SA) {
  // text holds symbols 1..4 followed by the sentinel 0
  SA.resize(text.size());
  sais<uint8_t, I>(text.data(), SA.data(), (I)text.size(), (I)5);
  return true;
}

template <typename I>
void fm_files_from_sa(const std::vector<uint8_t> &text, const std::vector<I> &SA, int dense_intv, FmFiles &f) {
  const uint64_t n = text.size() - 1;  // text length without sentinel (= 2 * l_pac)
  f.n = n;
  // BWT rows 0..n; row 0 is the suffix "$".  Occ counts exclude '$'.
  uint64_t cnt[4] = {0, 0, 0, 0};
  for (uint64_t i = 0; i < n; ++i) ++cnt[text[i] - 1];
  f.L2[0] = 0;
  for (int x = 0; x < 4; ++x) f.L2[x + 1] = f.L2[x] + cnt[x];
  // BWA packs the BWT without the '$' row: position k of the string is row
  // k (k < primary) or row k+1 (k >= primary).
  const uint64_t n_occ = (n + 127) / 128 + 1;
  const uint64_t raw_words = (n + 15) / 16;
  const uint64_t bwt_words = raw_words + n_occ * 8;
  f.bwt.assign((size_t)bwt_words, 0);
  uint64_t c[4] = {0, 0, 0, 0};
  uint64_t k = 0, w = 0;
  uint32_t cur = 0;
  for (uint64_t row = 0; row <= n; ++row) {
    uint64_t sa = (uint64_t)SA[(size_t)row];
    if (sa == 0) { f.primary = row; continue; }
    uint32_t b = (uint32_t)(text[(size_t)(sa - 1)] - 1);
    if (k % 128 == 0) {
      std::memcpy(&f.bwt[(size_t)w], c, sizeof c);
      w += 8;
    }
    cur |= b << ((15 - (k & 15)) << 1);
    if ((k & 15) == 15) { f.bwt[(size_t)w++] = cur; cur = 0; }
    ++c[b];
    ++k;
  }
  if (k & 15) f.bwt[(size_t)w++] = cur;
  std::memcpy(&f.bwt[(size_t)w], c, sizeof c);
  // BWA sampled SA: SA[32*i] for i in [1, n_sa)
  const uint64_t n_sa = (n + 32) / 32;
  f.sa.resize((size_t)(n_sa - 1));
  for (uint64_t i = 1; i < n_sa; ++i) f.sa[(size_t)(i - 1)] = (uint64_t)SA[(size_t)(i * 32)];
  // bwa-mem2 .counts: cumulative count[5], cumulative count2[17]
  int64_t cnt1[5] = {0, 0, 0, 0, 0}, cnt2[17];
  std::memset(cnt2, 0, sizeof cnt2);
  for (uint64_t i = 0; i < n; ++i) ++cnt1[text[i] - 1];
  for (uint64_t i = 0; i + 1 < n; ++i) ++cnt2[(text[i] - 1) * 4 + (text[i + 1] - 1)];
  int64_t acc = 0;
  for (int i = 0; i < 5; ++i) { f.cnt1[i] = acc; acc += cnt1[i]; }
  acc = 0;
  for (int i = 0; i < 17; ++i) { f.cnt2[i] = acc; acc += cnt2[i]; }
  // dense SA sample for the device: SA[intv*i] for i in [0, ceil((n+1)/intv))
  f.g3_intv = (uint32_t)dense_intv;
  f.g3_width = n + 1 < 0xffffffffULL ? 4 : 8;
  f.g3_count = (n + 1 + (uint64_t)dense_intv - 1) / (uint64_t)dense_intv;
  f.g3sa.resize((size_t)(f.g3_count * f.g3_width));
  for (uint64_t i = 0; i < f.g3_count; ++i) {
    const uint64_t v = (uint64_t)SA[(size_t)(i * f.g3_intv)];
    if (f.g3_width == 4) {
      const uint32_t v32 = (uint32_t)v;
      std::memcpy(&f.g3sa[(size_t)(i * 4)], &v32, 4);
    } else {
      std::memcpy(&f.g3sa[(size_t)(i * 8)], &v, 8);
    }
  }
}

}  // namespace

bool write_fm_files(const std::string &prefix, const FmFiles &f, std::string &err) {
  const uint64_t n = f.n;
  {
    std::string fn = prefix + ".bwt";
    FILE *fp = std::fopen(fn.c_str(), "wb");
    if (!fp) { err = "cannot write " + fn; return false; }
    bool ok = std::fwrite(&f.primary, 8, 1, fp) == 1 && std::fwrite(f.L2 + 1, 8, 4, fp) == 4 &&
              std::fwrite(f.bwt.data(), 4, f.bwt.size(), fp) == f.bwt.size();
    ok = std::fclose(fp) == 0 && ok;
    if (!ok) { err = "short write to " + fn; return false; }
  }
  {  // BWA .sa: header then SA[32*i] for i in [1, n_sa)
    const uint64_t intv = 32;
    std::string fn = prefix + ".sa";
    FILE *fp = std::fopen(fn.c_str(), "wb");
    if (!fp) { err = "cannot write " + fn; return false; }
    bool ok = std::fwrite(&f.primary, 8, 1, fp) == 1 && std::fwrite(f.L2 + 1, 8, 4, fp) == 4 &&
              std::fwrite(&intv, 8, 1, fp) == 1 && std::fwrite(&n, 8, 1, fp) == 1 &&
              std::fwrite(f.sa.data(), 8, f.sa.size(), fp) == f.sa.size();
    ok = std::fclose(fp) == 0 && ok;
    if (!ok) { err = "short write to " + fn; return false; }
  }
  {  // bwa-mem2 .counts: n+1, cumulative count[5], cumulative count2[17]
    const int64_t rows = (int64_t)n + 1;
    std::string fn = prefix + ".counts";
    FILE *fp = std::fopen(fn.c_str(), "wb");
    if (!fp) { err = "cannot write " + fn; return false; }
    bool ok = std::fwrite(&rows, 8, 1, fp) == 1 && std::fwrite(f.cnt1, 8, 5, fp) == 5 &&
              std::fwrite(f.cnt2, 8, 17, fp) == 17;
    ok = std::fclose(fp) == 0 && ok;
    if (!ok) { err = "short write to " + fn; return false; }
  }
  {
    G3SaHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, "G3SA0001", 8);
    h.rows = n + 1;
    h.intv = f.g3_intv;
    h.width = f.g3_width;
    h.count = f.g3_count;
    std::string fn = prefix + ".g3sa";
    FILE *fp = std::fopen(fn.c_str(), "wb");
    if (!fp) { err = "cannot write " + fn; return false; }
    bool ok = std::fwrite(&h, sizeof h, 1, fp) == 1 &&
              std::fwrite(f.g3sa.data(), 1, f.g3sa.size(), fp) == f.g3sa.size();
    ok = std::fclose(fp) == 0 && ok;
    if (!ok) { err = "short write to " + fn; return false; }
  }
  return true;
}

bool write_bns(const std::string &prefix, const std::vector<std::string> &names,
               const std::vector<std::string> &annos, const std::vector<std::vector<uint8_t>> &seqs,
               std::vector<uint8_t> &fwd, std::string &err) {
  // Ambiguous runs become holes (.amb) and are filled with generated bases.
  struct Hole { int64_t off, len; char c; };
  std::vector<Hole> holes;
  std::vector<int> n_amb(seqs.size(), 0);
  HoleRng rng(11);
  fwd.clear();
  int64_t l_pac = 0;
  for (auto &s : seqs) l_pac += (int64_t)s.size();
  fwd.reserve((size_t)l_pac);
  for (size_t i = 0; i < seqs.size(); ++i) {
    const auto &s = seqs[i];
    for (size_t j = 0; j < s.size(); ++j) {
      uint8_t b = s[j];
      if (b > 3) {
        int64_t off = (int64_t)fwd.size();
        if (!holes.empty() && holes.back().off + holes.back().len == off) ++holes.back().len;
        else { holes.push_back({off, 1, 'N'}); ++n_amb[i]; }
        b = (uint8_t)(rng.next() & 3);
      }
      fwd.push_back(b);
    }
  }
  {
    std::vector<uint8_t> pac((size_t)(l_pac / 4 + 2), 0);
    for (int64_t i = 0; i < l_pac; ++i) pac[(size_t)(i >> 2)] |= (uint8_t)(fwd[(size_t)i] << ((~i & 3) << 1));
    size_t bytes = (size_t)((l_pac >> 2) + ((l_pac & 3) ? 1 : 0));
    std::vector<uint8_t> out(pac.begin(), pac.begin() + (long)bytes);
    if (l_pac % 4 == 0) out.push_back(0);
    out.push_back((uint8_t)(l_pac % 4));
    if (!write_all(prefix + ".pac", out.data(), out.size(), err)) return false;
  }
  {
    std::string fn = prefix + ".ann";
    FILE *fp = std::fopen(fn.c_str(), "w");
    if (!fp) { err = "cannot write " + fn; return false; }
    std::fprintf(fp, "%lld %d %u\n", (long long)l_pac, (int)seqs.size(), 11u);
    int64_t off = 0;
    for (size_t i = 0; i < seqs.size(); ++i) {
      std::fprintf(fp, "0 %s", names[i].c_str());
      if (!annos[i].empty()) std::fprintf(fp, " %s\n", annos[i].c_str());
      else std::fprintf(fp, "\n");
      std::fprintf(fp, "%lld %lld %d\n", (long long)off, (long long)seqs[i].size(), n_amb[i]);
      off += (int64_t)seqs[i].size();
    }
    std::fclose(fp);
  }
  {
    std::string fn = prefix + ".amb";
    FILE *fp = std::fopen(fn.c_str(), "w");
    if (!fp) { err = "cannot write " + fn; return false; }
    std::fprintf(fp, "%lld %d %u\n", (long long)l_pac, (int)seqs.size(), (unsigned)holes.size());
    for (auto &h : holes) std::fprintf(fp, "%lld %lld %c\n", (long long)h.off, (long long)h.len, h.c);
    std::fclose(fp);
  }
  return true;
}

bool build_fm_from_fwd(const std::string &prefix, const std::vector<uint8_t> &fwd, int dense_intv,
                       std::string &err, int device) {
  const uint64_t L = fwd.size();
  if (L == 0) { err = "empty reference"; return false; }
  FmFiles f;
  if (device >= 0) {
    if (!build_fm_device(device, fwd, dense_intv, f, err)) return false;
  } else {
    std::vector<uint8_t> text((size_t)(2 * L + 1));
    for (uint64_t i = 0; i < L; ++i) {
      text[(size_t)i] = (uint8_t)(fwd[(size_t)i] + 1);
      text[(size_t)(2 * L - 1 - i)] = (uint8_t)(4 - fwd[(size_t)i]);  // complement, symbol = (3-b)+1
    }
    text[(size_t)(2 * L)] = 0;
    if (text.size() < 0xffffffffULL) {
      std::vector<uint32_t> SA;
      build_sa(text, SA);
      fm_files_from_sa(text, SA, dense_intv, f);
    } else {
      std::vector<uint64_t> SA;
      build_sa(text, SA);
      fm_files_from_sa(text, SA, dense_intv, f);
    }
  }
  const bool ok = write_fm_files(prefix, f, err);
  if (ok && sais_progress) sais_progress(".bwt .sa .g3sa .counts written", 0, (int64_t)(2 * L + 1));
  return ok;
}

void set_build_progress(void (*fn)(const char *what, int level, int64_t n)) { sais_progress = fn; }

bool build_index(const char *fasta, const char *prefix, int dense_intv, std::string &err, int device) {
  if (dense_intv < 1 || dense_intv > 32 || (dense_intv & (dense_intv - 1))) {
    err = "SA sample interval must be a power of two in [1, 32]";
    return false;
  }
  std::vector<FaContig> ctg;
  if (!read_fasta(fasta, ctg, err)) return false;
  std::vector<std::string> names, annos;
  std::vector<std::vector<uint8_t>> seqs;
  for (auto &c : ctg) {
    names.push_back(c.name);
    annos.push_back(c.anno);
    seqs.push_back(std::move(c.seq));
  }
  if (sais_progress) sais_progress("reference read", 0, 0);
  std::vector<uint8_t> fwd;
  if (!write_bns(prefix, names, annos, seqs, fwd, err)) return false;
  if (sais_progress) sais_progress(".pac .ann .amb written", 0, (int64_t)fwd.size());
  return build_fm_from_fwd(prefix, fwd, dense_intv, err, device);
}

}  // namespace g3a
