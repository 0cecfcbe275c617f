// Output records of one batch, formatted per read (host or device): the
// reference's record lines (bwa/src/pipeline.cpp:287-351) or standard SAM
// records, plus MAPQ (mem_approx_mapq_se, bwa/src/aux.cu:676-700) and the
// MD string (aux.cu:2908-2941 semantics, recomputed in genome orientation).
//
// emit_read() is written once against a sink: a counting sink gives each
// read's byte length (pass 1), a scan turns the lengths into offsets, and a
// writing sink stores the bytes (pass 2), so the two passes cannot disagree.
// The writing sink gathers bytes in a 16-byte register accumulator and
// stores whole aligned 16-byte chunks (the head and tail of a read's text,
// which share chunks with its neighbours, go out byte by byte).  The bulk of
// a record, SEQ and QUAL, is not written by the thread of its read: the
// counting pass records where each record's SEQ goes (a TextSlot), the
// writing pass skips those bytes, and k_text_fill writes them with one
// wavefront per record, its lanes on consecutive bytes.
#pragma once
#include "g3a_dev.h"

#include <cmath>

namespace g3a {

// logf(l) for the MAPQ length term, tabulated on the host with the C
// library, so that the device's MAPQ equals the host's bit for bit.
constexpr int kLogTab = 4096;

// Where a record's SEQ starts in its read's text (QUAL follows SEQ and one
// separator), and which way it runs.  Read q's records own slots
// aln_off[q] + q .. aln_off[q + 1] + q: one per alignment, or one for an
// unmapped read; a mapped read's last slot is unused (read = -1).
struct TextSlot {
  int64_t seq_at;  // offset in the read's text
  int32_t read;    // -1: unused
  int32_t rev;     // reverse-complemented SEQ and reversed QUAL
};

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
  int32_t *nm;  // per alignment: SAM NM, written by the counting pass, read by the writing pass
  TextSlot *slots;  // n_alns + n_reads, written by the counting pass
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
  static constexpr bool kCount = true;
  int64_t n = 0;
  G3A_HD void skip(int64_t len) { n += len; }
  G3A_HD void put(char) { ++n; }
  G3A_HD void put(const char *s, int64_t len) { n += len; }
  G3A_HD void lit(const char *s) {
    while (*s++) ++n;
  }
};

// Plain byte stores (the MD strings of the output arrays).
struct WriteSink {
  static constexpr bool kCount = false;
  char *p;
  G3A_HD void put(char c) { *p++ = c; }
  G3A_HD void put(const char *s, int64_t len) {
    for (int64_t i = 0; i < len; ++i) p[i] = s[i];
    p += len;
  }
  G3A_HD void lit(const char *s) {
    while (*s) *p++ = *s++;
  }
  G3A_HD void flush() {}
};

struct alignas(16) Chunk16 {
  uint64_t lo, hi;
};

// Byte stores until the address is 16-aligned, then bytes gathered into a
// 16-byte register accumulator stored as one aligned chunk; flush() writes
// the last partial chunk byte by byte.
struct WriteSink16 {
  static constexpr bool kCount = false;
  char *p;            // next byte's address
  uint64_t lo = 0, hi = 0;
  int fill = 0;       // bytes in the accumulator (p is 16-aligned when fill > 0 or after the head)
  bool head = true;   // still before the first 16-byte boundary
  G3A_HD explicit WriteSink16(char *dst) : p(dst) { head = ((uintptr_t)p & 15) != 0; }
  G3A_HD void put(char c) {
    if (head) {
      *p++ = c;
      head = ((uintptr_t)p & 15) != 0;
      return;
    }
    const uint64_t b = (uint64_t)(uint8_t)c;
    if (fill < 8) lo |= b << (8 * fill);
    else hi |= b << (8 * (fill - 8));
    if (++fill == 16) {
      *(Chunk16 *)p = Chunk16{lo, hi};
      p += 16;
      lo = hi = 0;
      fill = 0;
    }
  }
  G3A_HD void put(const char *s, int64_t len) {
    for (int64_t i = 0; i < len; ++i) put(s[i]);
  }
  G3A_HD void lit(const char *s) {
    while (*s) put(*s++);
  }
  G3A_HD void flush() {
    for (int i = 0; i < fill; ++i) p[i] = (char)((i < 8 ? lo >> (8 * i) : hi >> (8 * (i - 8))) & 0xff);
    p += fill;
    fill = 0;
    lo = hi = 0;
  }
  // leave len bytes to another writer: no chunk store may cover them
  G3A_HD void skip(int64_t len) {
    flush();
    p += len;
    head = ((uintptr_t)p & 15) != 0;
  }
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

// SEQ, the separator and QUAL of one record in slot `slot`: counted (and
// the slot filled) by the counting pass, skipped by the writing pass except
// for the separator; k_text_fill writes the bases and qualities.
template <class S>
G3A_HD void seq_qual(S &s, const TextArgs &t, int q, int64_t slot, int len, int64_t ql, bool rev, char sep) {
  if constexpr (S::kCount) t.slots[slot] = TextSlot{s.n, q, rev ? 1 : 0};
  s.skip(len);
  s.put(sep);
  s.skip(ql);
}

// All lines of read q.
template <class S>
G3A_HD void emit_read(S &s, const TextArgs &t, int q) {
  const char *kOp = "MIDSH";
  const char *name = t.names + t.name_off[q];
  const int64_t nl = t.name_off[q + 1] - t.name_off[q];
  const int64_t ql = t.qual_off[q + 1] - t.qual_off[q];
  const uint8_t *sq = t.seq + t.off[q];
  const int len = t.off[q + 1] - t.off[q];
  const int64_t a0 = t.aln_off[q], a1 = t.aln_off[q + 1];
  if constexpr (S::kCount) {
    if (a1 > a0) t.slots[a1 + q].read = -1;  // the mapped read's unused slot
  }
  if (t.mode == 0) {
    if (a1 <= a0) {
      s.put(name, nl);
      s.lit(" * 0 * ");
      seq_qual(s, t, q, a0 + q, len, ql, false, ' ');
      s.put('\n');
      return;
    }
    for (int64_t i = a0; i < a1; ++i) {
      const Aln &a = t.alns[i];
      s.put(name, nl);
      s.put(' ');
      s.put(t.ctg_names + t.ctg_name_off[a.rid], t.ctg_name_off[a.rid + 1] - t.ctg_name_off[a.rid]);
      s.put(' ');
      put_num(s, a.pos);
      s.put(' ');
      for (int k = 0; k < a.n_cigar; ++k) {
        const uint32_t c = t.cigar[a.cigar_off + k];
        put_num(s, c >> 4);
        s.put(kOp[c & 0xf]);
      }
      s.put(' ');
      seq_qual(s, t, q, i + q, len, ql, false, ' ');
      s.put('\n');
    }
    return;
  }
  if (a1 <= a0) {
    s.put(name, nl);
    s.lit("\t4\t*\t0\t0\t*\t*\t0\t0\t");
    seq_qual(s, t, q, a0 + q, len, ql, false, '\t');
    s.put('\n');
    return;
  }
  for (int64_t i = a0; i < a1; ++i) {
    const Aln &a = t.alns[i];
    const bool rev = a.rb >= t.l_pac;
    int64_t fpos;
    int beg, endc;
    sam_place(t, a, rev, fpos, beg, endc);
    const int clip_l = rev ? len - a.qe : a.qb, clip_r = rev ? a.qb : len - a.qe;
    const int flag = (rev ? 0x10 : 0) | (i > a0 ? 0x800 : 0);
    s.put(name, nl);
    s.put('\t');
    put_num(s, flag);
    s.put('\t');
    s.put(t.ctg_names + t.ctg_name_off[a.rid], t.ctg_name_off[a.rid + 1] - t.ctg_name_off[a.rid]);
    s.put('\t');
    put_num(s, fpos - t.ctg_off[a.rid] + 1);
    s.put('\t');
    put_num(s, mapq_se(t, a.score, a.qb, a.qe, a.rb, a.re));
    s.put('\t');
    if (clip_l) {
      put_num(s, clip_l);
      s.put('S');
    }
    for (int k = beg; k < endc; ++k) {
      const uint32_t c = core_op(t, a, rev, k);
      put_num(s, c >> 4);
      s.put(kOp[c & 0xf]);
    }
    if (clip_r) {
      put_num(s, clip_r);
      s.put('S');
    }
    s.lit("\t*\t0\t0\t");
    seq_qual(s, t, q, i + q, len, ql, rev, '\t');
    // NM precedes MD in the record, and both come from one walk: the
    // counting pass walks it, keeps NM for the writing pass and counts the
    // MD string; the writing pass writes NM, then walks to write the string
    if constexpr (S::kCount) {
      CountSink mdc;
      const int nm = emit_md(mdc, t, a, rev, sq, beg, endc, fpos);
      t.nm[i] = nm;
      s.lit("\tNM:i:");
      put_num(s, nm);
      s.lit("\tMD:Z:");
      s.put(nullptr, mdc.n);
    } else {
      s.lit("\tNM:i:");
      put_num(s, t.nm[i]);
      s.lit("\tMD:Z:");
      emit_md(s, t, a, rev, sq, beg, endc, fpos);
    }
    s.lit("\tAS:i:");
    put_num(s, a.score);
    s.put('\n');
  }
}

}  // namespace g3a
