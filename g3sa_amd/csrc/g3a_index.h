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
  std::vector<uint32_t> bwt_words;  // the .bwt body (BWA blocks of 128 positions)
  std::vector<uint8_t> bwt;     // rows 0..n as codes 0..3, 4 for '$' at `primary`
                                // (filled by unpack_bwt; load_index leaves it empty)
  uint32_t sa_intv = 0, sa_width = 0;
  std::vector<uint8_t> sa;      // sa_width-byte entries SA[i * sa_intv]
};

// The FM-index files of one text, as written to disk: BWA .bwt body words,
// BWA's 32-row SA sample, bwa-mem2 .counts tables and the dense .g3sa sample.
struct FmFiles {
  uint64_t n = 0, primary = 0;
  uint64_t L2[5] = {0, 0, 0, 0, 0};
  std::vector<uint32_t> bwt;  // .bwt after its 40-byte header
  std::vector<uint64_t> sa;   // SA[32 i] for i in [1, (n + 32) / 32)
  int64_t cnt1[5] = {0, 0, 0, 0, 0}, cnt2[17] = {};  // cumulative counts
  uint32_t g3_intv = 4, g3_width = 4;
  uint64_t g3_count = 0;
  std::vector<uint8_t> g3sa;  // g3_count entries of g3_width bytes
};
bool write_fm_files(const std::string &prefix, const FmFiles &f, std::string &err);
// Builds the files of fwd + revcomp on HIP device `device` (g3a_index_gpu.hip).
bool build_fm_device(int device, const std::vector<uint8_t> &fwd, int dense_intv, FmFiles &out, std::string &err);
// Progress lines of build_fm_device (phase, doubling round, suffixes left).
extern void (*device_build_progress)(const char *what, int level, int64_t n);

// device < 0: suffix array by SA-IS on the host; otherwise on that HIP device.
bool build_index(const char *fasta, const char *prefix, int dense_intv, std::string &err, int device = -1);
// Progress lines of build_index (phase, recursion level of the suffix sort,
// length sorted) go to fn; null (the default) keeps the build silent.
void set_build_progress(void (*fn)(const char *what, int level, int64_t n));
bool write_bns(const std::string &prefix, const std::vector<std::string> &names,
               const std::vector<std::string> &annos, const std::vector<std::vector<uint8_t>> &seqs,
               std::vector<uint8_t> &fwd, std::string &err);
bool build_fm_from_fwd(const std::string &prefix, const std::vector<uint8_t> &fwd, int dense_intv,
                       std::string &err, int device = -1);
bool load_index(const char *prefix, HostIndex &idx, std::string &err);
// Fills idx.bwt (one byte per row) from idx.bwt_words.
void unpack_bwt(HostIndex &idx);

}  // namespace g3a
