// Host-side index structures for the BWA-MEM hot path.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace g3a {

// Header of <prefix>.g3sa, the dense suffix-array sample: entry i holds
// SA[i * intv] as a `width`-byte unsigned integer.
struct G3SaHeader {
  char magic[8];      // "G3SA0001"
  uint64_t rows;      // n + 1 rows of the suffix array (n = 2 * l_pac)
  uint32_t intv;      // sampling interval (power of two)
  uint32_t width;     // 4 or 8
  uint64_t count;     // number of entries
  uint64_t reserved[3];
};

struct Contig {
  std::string name, anno;
  int64_t offset = 0, len = 0;
  int n_ambs = 0;
  int is_alt = 0;
};

// Everything the device pipeline needs, as loaded from disk.
struct HostIndex {
  int64_t l_pac = 0;            // forward reference length
  std::vector<Contig> anns;
  std::vector<uint8_t> pac;     // forward strand, 2 bits per base, MSB first
  uint64_t seq_len = 0;         // n = 2 * l_pac
  uint64_t primary = 0;         // BWT row whose suffix starts at text position 0
  uint64_t L2[5] = {0, 0, 0, 0, 0};  // cumulative base counts (no '$')
  std::vector<uint8_t> bwt;     // rows 0..n as codes 0..3, 4 for '$' at `primary`
  uint32_t sa_intv = 0, sa_width = 0;
  std::vector<uint8_t> sa;      // sa_width-byte entries SA[i * sa_intv]
};

bool build_index(const char *fasta, const char *prefix, int dense_intv, std::string &err);
// Progress lines of build_index (phase, recursion level of the suffix sort,
// length sorted) go to fn; null (the default) keeps the build silent.
void set_build_progress(void (*fn)(const char *what, int level, int64_t n));
bool write_bns(const std::string &prefix, const std::vector<std::string> &names,
               const std::vector<std::string> &annos, const std::vector<std::vector<uint8_t>> &seqs,
               std::vector<uint8_t> &fwd, std::string &err);
bool build_fm_from_fwd(const std::string &prefix, const std::vector<uint8_t> &fwd, int dense_intv,
                       std::string &err);
bool load_index(const char *prefix, HostIndex &idx, std::string &err);

}  // namespace g3a
