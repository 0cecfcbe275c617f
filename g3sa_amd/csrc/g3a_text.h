// Output records of one batch, formatted on the device: the reference's
// record lines (bwa/src/pipeline.cpp:287-351) or standard SAM records, plus
// MAPQ (mem_approx_mapq_se, bwa/src/aux.cu:676-700) and the MD string
// (aux.cu:2908-2941 semantics, recomputed in genome orientation).  This
// header holds the per-record pieces shared by the record formatter and the
// C ABI's MAPQ/MD arrays (g3a_text.hip); emit_md is written once against a
// sink, so that a counting sink gives the MD string's length and a writing
// sink its bytes.
#pragma once
#include "g3a_dev.h"

#include <cmath>

namespace g3a {

// logf(l) for the MAPQ length term, tabulated on the host with the C
// library, so that the device's MAPQ equals the host's bit for bit.
constexpr int kLogTab = 4096;

struct TextArgs {
  int mode;  // 0: the reference's record lines; 1: SAM records
  int n_reads;
  const uint8_t *seq;     // 2-bit codes (4 = N)
  const int32_t *off;     // n_reads + 1
  const char *names;
  const int64_t *name_off;  // n_reads + 1
  const char *quals;
  const int64_t *qual_off;  // n_reads + 1
  const char *ctg_names;
  const int64_t *ctg_name_off;  // n_contigs + 1
  const int64_t *ctg_off;       // contig start in the forward text
  const uint8_t *pac;           // forward strand, 2 bits per base, MSB first
  int64_t l_pac;
  const int64_t *aln_off;  // n_reads + 1
  const Aln *alns;
  const uint32_t *cigar;
  const float *log_tab;  // kLogTab entries: logf((float)i)
  int a, b, min_seed_len;
};

// mem_approx_mapq_se with sub = min_seed_len * a (no sub-optimal hits are
// recorded on this path), csub = sub_n = 0 and frac_rep = 0 (DR-3)
G3A_HD int mapq_se(const TextArgs &t, int score, int qb, int qe, int64_t rb, int64_t re) {
  const int sub = t.min_seed_len * t.a;
  if (sub >= score) return 0;
  const int l = qe - qb > re - rb ? qe - qb : (int)(re - rb);
  const double identity = 1. - (double)(l * t.a - score) / (t.a + t.b) / l;
  int mapq;
  if (score == 0) {
    mapq = 0;
  } else {
    const int coef_len = 50, coef_fac = 3;  // mapQ_coef_len, (int)log(50)
    const float lg = l < kLogTab ? t.log_tab[l] : logf((float)l);
    double tmp = l < coef_len ? 1. : coef_fac / lg;
    tmp *= identity * identity;
    mapq = (int)(6.02 * (score - sub) / t.a * tmp * tmp + .499);
  }
  if (mapq > 60) mapq = 60;
  if (mapq < 0) mapq = 0;
  return mapq;
}

G3A_HD int ref_base(const uint8_t *pac, int64_t l) { return (pac[l >> 2] >> ((~l & 3) << 1)) & 3; }

struct CountSink {
  int64_t n = 0;
  G3A_HD void put(char) { ++n; }
};

// plain byte stores
struct WriteSink {
  char *p;
  G3A_HD void put(char c) { *p++ = c; }
};

template <class S>
G3A_HD void put_num(S &s, int64_t v) {
  char t[24];
  int k = 0;
  uint64_t u = v < 0 ? 0ull - (uint64_t)v : (uint64_t)v;
  do {
    t[k++] = (char)('0' + u % 10);
    u /= 10;
  } while (u);
  if (v < 0) s.put('-');
  while (k) s.put(t[--k]);
}

// core CIGAR op k of alignment a in genome orientation (reversed for the
// reverse strand)
G3A_HD uint32_t core_op(const TextArgs &t, const Aln &a, bool rev, int k) {
  return t.cigar[a.core_off + (rev ? a.n_core - 1 - k : k)];
}

// One SAM record's MD string and NM; when S is a counting sink only the
// length matters.  beg/endc: the core ops without a leading or trailing D.
template <class S>
G3A_HD int emit_md(S &s, const TextArgs &t, const Aln &a, bool rev, const uint8_t *sq, int beg, int endc,
                   int64_t fpos) {
  int nm = 0, run = 0, xq = 0;
  int64_t yr = fpos;
  for (int k = beg; k < endc; ++k) {
    const uint32_t c = core_op(t, a, rev, k);
    const int op = c & 0xf, ln = (int)(c >> 4);
    if (op == 0) {
      for (int u = 0; u < ln; ++u) {
        const int qi = xq + u;
        uint8_t b = rev ? sq[a.qe - 1 - qi] : sq[a.qb + qi];
        if (rev && b < 4) b = (uint8_t)(3 - b);
        const int rbase = ref_base(t.pac, yr + u);
        if (b != rbase) {
          put_num(s, run);
          s.put("ACGT"[rbase]);
          run = 0;
          ++nm;
        } else {
          ++run;
        }
      }
      xq += ln;
      yr += ln;
    } else if (op == 1) {
      xq += ln;
      nm += ln;
    } else if (op == 2) {
      put_num(s, run);
      s.put('^');
      for (int u = 0; u < ln; ++u) s.put("ACGT"[ref_base(t.pac, yr + u)]);
      run = 0;
      yr += ln;
      nm += ln;
    }
  }
  put_num(s, run);
  return nm;
}

// Genome-oriented placement of alignment a: leading-D-adjusted forward
// position and the core op range [beg, endc).
G3A_HD void sam_place(const TextArgs &t, const Aln &a, bool rev, int64_t &fpos, int &beg, int &endc) {
  fpos = rev ? (t.l_pac << 1) - a.re : a.rb;
  beg = 0;
  endc = a.n_core;
  if (endc > 0 && (core_op(t, a, rev, 0) & 0xf) == 2) {
    fpos += core_op(t, a, rev, 0) >> 4;
    beg = 1;
  }
  if (endc > beg && (core_op(t, a, rev, endc - 1) & 0xf) == 2) --endc;
}

}  // namespace g3a
