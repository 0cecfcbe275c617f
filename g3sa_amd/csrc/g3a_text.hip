// Device formatting of a batch's output records, one wavefront per read,
// and each alignment's MAPQ and MD string for the C ABI's output arrays.
//
// A record is NAME, a run of head tokens, SEQ, a separator, QUAL and a run
// of tail tokens.  A token is a literal, a decimal number, one character and
// a byte string, each part optional: a field such as "\tPOS", one CIGAR op,
// one MD event ("12A", "3^AC").  The 64 lanes of the read's wavefront take
// 64 tokens at a time: each lane selects its token's parts (cheap, divergent
// selects), then all lanes measure and write their tokens with the same
// code, so the formatting itself runs 64 wide; a wave prefix sum over the
// token lengths places them.  NAME, SEQ and QUAL are copied by all lanes on
// consecutive bytes.  The MD events come from a walk over the aligned bases,
// 64 at a time: a ballot gives the chunk's mismatches, and each mismatching
// lane writes its own event (its run length and reference base) into the
// wavefront's LDS list.  The counting pass stores each read's text length, a
// scan turns the lengths into offsets, and the writing pass (the same code)
// stores the text, so the two passes cannot disagree.
//
// Record layout (the reference's lines: bwa/src/pipeline.cpp:287-351; SAM
// records with MAPQ, NM, MD and AS):
//          SAM mapped             SAM unmapped                record lines
//   NAME   NAME                   NAME                        NAME
//   head   \tFLAG \tRNAME \tPOS    \t4\t*\t0\t0\t*\t*\t0\t0\t  ' 'RNAME ' 'POS ' '
//          \tMAPQ \t[clip S]                                  CIGAR ops ' '
//          CIGAR ops [clip S]                                 (unmapped: " * 0 * ")
//          \t*\t0\t0\t
//   SEQ    SEQ \t QUAL            SEQ \t QUAL                 SEQ ' ' QUAL
//   tail   \tNM:i:NM \tMD:Z:       \n                          \n
//          MD events, final run
//          \tAS:i:AS\n
#include "g3a_kernels.h"
#include "g3a_text.h"

#include <hip/hip_runtime.h>

namespace g3a {

namespace {

constexpr int kFmtWaves = 4;    // reads (wavefronts) per block
constexpr int kMdEvents = 256;  // MD events held in LDS per record at a time (a window of its list)
constexpr int kRegBytes = 4;    // bytes per lane of SEQ, NAME and QUAL held in registers (fields up to 256 bytes)

struct MdEvent {
  uint32_t run;   // matching bases before the event
  uint32_t yr;    // deletion: its first reference base, from the record's first
  uint32_t dlen;  // deletion: its length
  uint8_t kind;   // 0 mismatch, 1 deletion, 2 the final run
  uint8_t base;   // mismatch: the reference base
};

// Letters by code, packed into constants (a register shift, not a table
// load): byte c of kAcgtn is the base of code c (4 = N), byte op of kOps the
// CIGAR letter of op.
constexpr uint64_t kAcgtn = 0x4e54474341ull;  // "ACGTN"
constexpr uint64_t kOps = 0x485344494dull;    // "MIDSH"
__device__ __forceinline__ char base_char(int c) { return (char)(kAcgtn >> (8 * (c < 4 ? c : 4))); }
__device__ __forceinline__ char op_char(int op) { return (char)(kOps >> (8 * (op < 5 ? op : 4))); }

// One token of a record: a literal of up to 16 bytes (pre0, pre1, packed
// little-endian), then a decimal number (has_num), then one character (c,
// when nonzero), then suf_len bytes: from suf, or, with suf null, reference
// bases from position ref on.
struct Tok {
  uint64_t pre0, pre1;
  const char *suf;
  int64_t num;
  int64_t ref;
  int pre_len, suf_len;
  char c;
  bool has_num;
};

// The fields of one record that every lane holds.  Events ev_from ..
// ev_from + kMdEvents - 1 of its MD list are in the wavefront's LDS window.
struct Rec {
  Aln a;
  bool mapped, sam, rev;
  int flag, beg, endc, clip_l, clip_r, n_ev, ev_from;
  int64_t fpos;
};

template <int N>
constexpr uint64_t pack8(const char (&s)[N], int from) {
  uint64_t v = 0;
  for (int k = 0; k < 8 && from + k < N - 1; ++k) v |= (uint64_t)(uint8_t)s[from + k] << (8 * k);
  return v;
}

template <int N>
__device__ __forceinline__ void set_pre(Tok &d, const char (&s)[N]) {
  static_assert(N - 1 <= 16, "token literals hold at most 16 bytes");
  d.pre0 = pack8(s, 0);
  d.pre1 = pack8(s, 8);
  d.pre_len = N - 1;
}

__device__ __forceinline__ Tok empty_tok() { return Tok{0, 0, nullptr, 0, 0, 0, 0, 0, false}; }

__device__ __forceinline__ int ndig_u(uint64_t u) {
  if (u >> 32) {
    int d = 10;
    u /= 10000000000ull;
    while (u) {
      u /= 10;
      ++d;
    }
    return d;
  }
  const uint32_t v = (uint32_t)u;
  return 1 + (v >= 10u) + (v >= 100u) + (v >= 1000u) + (v >= 10000u) + (v >= 100000u) + (v >= 1000000u) +
         (v >= 10000000u) + (v >= 100000000u) + (v >= 1000000000u);
}

__device__ __forceinline__ int ndig(int64_t v) {
  return v < 0 ? 1 + ndig_u(0ull - (uint64_t)v) : ndig_u((uint64_t)v);
}

__device__ __forceinline__ char *put_dec(char *p, int64_t v) {
  uint64_t u = v < 0 ? 0ull - (uint64_t)v : (uint64_t)v;
  if (v < 0) *p++ = '-';
  const int d = ndig_u(u);
  int k = d - 1;
  while (u >> 32) {
    p[k--] = (char)('0' + u % 10);
    u /= 10;
  }
  uint32_t w = (uint32_t)u;
  for (; k >= 0; --k) {
    p[k] = (char)('0' + w % 10u);
    w /= 10u;
  }
  return p + d;
}

// exclusive prefix sum of v over the wavefront's 64 lanes
__device__ __forceinline__ int wave_excl_scan(int v, int lane) {
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x - v;
}

// byte k of a lane's packed register (byte k holds field byte lane + 64k)
__device__ __forceinline__ int reg_byte(uint32_t w, int k) { return (int)((w >> (8 * k)) & 0xff); }

// code of base i of the read (i < len): from the lanes' registers when the
// read fits them, else from memory
__device__ __forceinline__ int read_code(uint32_t rs, const uint8_t *sq, int len, int i) {
  if (len <= 64 * kRegBytes) return reg_byte((uint32_t)__shfl((int)rs, i & 63), i >> 6);
  return sq[i];
}

// NM of SAM record r and its MD events (the mismatches found 64 aligned
// bases at a time, one per lane, each mismatching lane placing its own
// event), *n_ev their number (the final run not counted); events ev_from ..
// ev_from + kMdEvents - 1 of the list, the final run included, go to ev.  The
// events give the string emit_md writes.  Two chunks are loaded before either
// is compared, so that their reference reads overlap.
__device__ int md_walk(const TextArgs &t, const Rec &r, uint32_t rs, const uint8_t *sq, int len, int lane,
                       MdEvent *ev, int ev_from, int *n_ev) {
  const Aln &a = r.a;
  int nm = 0, xq = 0, nev = 0, run = 0;
  int64_t yr = 0;
  for (int k = r.beg; k < r.endc; ++k) {
    const uint32_t c = core_op(t, a, r.rev, k);
    const int op = c & 0xf, ln = (int)(c >> 4);
    if (op == 0) {
      for (int c0 = 0; c0 < ln; c0 += 128) {
        int rb[2], qb[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int u = c0 + 64 * h + lane;
          rb[h] = u < ln ? ref_base(t.pac, r.fpos + yr + u) : 0;
          const int qi = xq + (u < ln ? u : 0);
          const int code = read_code(rs, sq, len, r.rev ? a.qe - 1 - qi : a.qb + qi);
          qb[h] = r.rev && code < 4 ? 3 - code : code;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int cc = c0 + 64 * h;
          if (cc >= ln) break;
          const bool mis = cc + lane < ln && qb[h] != rb[h];
          const uint64_t m = __ballot(mis);
          if (mis) {
            const uint64_t low = m & ((1ull << lane) - 1);
            const int idx = nev + __popcll(low) - ev_from;
            const int rr = low ? lane - (63 - __clzll(low)) - 1 : run + lane;
            if (idx >= 0 && idx < kMdEvents) ev[idx] = MdEvent{(uint32_t)rr, 0, 0, 0, (uint8_t)rb[h]};
          }
          const int clen = ln - cc < 64 ? ln - cc : 64;
          run = m ? clen - 1 - (63 - __clzll(m)) : run + clen;
          nev += __popcll(m);
          nm += __popcll(m);
        }
      }
      xq += ln;
      yr += ln;
    } else if (op == 1) {
      xq += ln;
      nm += ln;
    } else if (op == 2) {
      if (lane == 0 && nev >= ev_from && nev - ev_from < kMdEvents)
        ev[nev - ev_from] = MdEvent{(uint32_t)run, (uint32_t)yr, (uint32_t)ln, 1, 0};
      ++nev;
      run = 0;
      yr += ln;
      nm += ln;
    }
  }
  if (lane == 0 && nev >= ev_from && nev - ev_from < kMdEvents) ev[nev - ev_from] = MdEvent{(uint32_t)run, 0, 0, 2, 0};
  __builtin_amdgcn_wave_barrier();
  *n_ev = nev;
  return nm;
}

// the tokens between NAME and SEQ
__device__ __forceinline__ int head_count(const Rec &r) {
  if (!r.mapped) return r.sam ? 2 : 1;
  return r.sam ? 7 + (r.endc - r.beg) : 4 + r.a.n_cigar;
}

__device__ Tok head_tok(const TextArgs &t, const Rec &r, int tk) {
  Tok d = empty_tok();
  const Aln &a = r.a;
  if (!r.mapped) {
    if (!r.sam)
      set_pre(d, " * 0 * ");
    else if (tk == 0)
      set_pre(d, "\t4\t*\t0\t0\t*\t*\t");
    else
      set_pre(d, "0\t0\t");
    return d;
  }
  d.pre0 = r.sam ? '\t' : ' ';
  d.pre_len = 1;
  if (r.sam) {
    const int nops = r.endc - r.beg;
    if (tk == 0) {
      d.has_num = true;
      d.num = r.flag;
    } else if (tk == 1) {
      d.suf = t.ctg_names + t.ctg_name_off[a.rid];
      d.suf_len = (int)(t.ctg_name_off[a.rid + 1] - t.ctg_name_off[a.rid]);
    } else if (tk == 2) {
      d.has_num = true;
      d.num = r.fpos - t.ctg_off[a.rid] + 1;
    } else if (tk == 3) {
      d.has_num = true;
      d.num = mapq_se(t, a.score, a.qb, a.qe, a.rb, a.re);
    } else if (tk == 4) {
      d.has_num = r.clip_l != 0;
      d.num = r.clip_l;
      d.c = r.clip_l ? 'S' : 0;
    } else if (tk < 5 + nops) {
      const uint32_t c = core_op(t, a, r.rev, r.beg + tk - 5);
      d.pre_len = 0;
      d.has_num = true;
      d.num = c >> 4;
      d.c = op_char(c & 0xf);
    } else if (tk == 5 + nops) {
      d.pre_len = 0;
      d.has_num = r.clip_r != 0;
      d.num = r.clip_r;
      d.c = r.clip_r ? 'S' : 0;
    } else {
      set_pre(d, "\t*\t0\t0\t");
    }
  } else {
    if (tk == 0) {
      d.suf = t.ctg_names + t.ctg_name_off[a.rid];
      d.suf_len = (int)(t.ctg_name_off[a.rid + 1] - t.ctg_name_off[a.rid]);
    } else if (tk == 1) {
      d.has_num = true;
      d.num = a.pos;
    } else if (tk >= 3 && tk < 3 + a.n_cigar) {
      const uint32_t c = t.cigar[a.cigar_off + tk - 3];
      d.pre_len = 0;
      d.has_num = true;
      d.num = c >> 4;
      d.c = op_char(c & 0xf);
    }
  }
  return d;
}

// the tokens after QUAL
__device__ __forceinline__ int tail_count(const Rec &r) {
  if (!r.mapped || !r.sam) return 1;
  return r.n_ev + 4;
}

__device__ Tok tail_tok(const Rec &r, int nm, const MdEvent *ev, int tk) {
  Tok d = empty_tok();
  if (!r.mapped || !r.sam) {
    d.c = '\n';
    return d;
  }
  const int last = tail_count(r) - 1;
  if (tk == 0) {
    set_pre(d, "\tNM:i:");
    d.has_num = true;
    d.num = nm;
  } else if (tk == last) {
    set_pre(d, "\tAS:i:");
    d.has_num = true;
    d.num = r.a.score;
    d.c = '\n';
  } else if (tk == 1) {
    set_pre(d, "\tMD:Z:");
  } else {
    const MdEvent E = ev[tk - 2 - r.ev_from];
    d.has_num = true;
    d.num = E.run;
    if (E.kind == 0) {
      d.c = base_char(E.base);
    } else if (E.kind == 1) {
      d.c = '^';
      d.ref = r.fpos + E.yr;
      d.suf_len = (int)E.dlen;
    }
  }
  return d;
}

__device__ __forceinline__ int tok_len(const Tok &d) {
  return d.pre_len + (d.has_num ? ndig(d.num) : 0) + (d.c ? 1 : 0) + d.suf_len;
}

// A token's bytes at p.
__device__ __forceinline__ void tok_write(const TextArgs &t, const Tok &d, char *p) {
  uint64_t lo = d.pre0, hi = d.pre1;
  for (int k = 0; k < d.pre_len; ++k) {
    p[k] = (char)(lo & 0xff);
    lo = (lo >> 8) | (hi << 56);
    hi >>= 8;
  }
  p += d.pre_len;
  if (d.has_num) p = put_dec(p, d.num);
  if (d.c) *p++ = d.c;
  for (int k = 0; k < d.suf_len; ++k) p[k] = d.suf ? d.suf[k] : base_char(ref_base(t.pac, d.ref + k));
}

// The head (kHead) or tail tokens of record r, 64 at a time, one per lane:
// their lengths, a wave scan for their places, and (kWrite) their bytes at
// out + at.  Returns their total length.  A round of tail tokens whose MD
// events lie outside the LDS window walks the alignment again for them.
template <bool kWrite, bool kHead>
__device__ int64_t emit_tokens(const TextArgs &t, Rec &r, uint32_t rs, const uint8_t *sq, int len, int nm,
                               MdEvent *ev, int lane, char *out, int64_t at) {
  const int T = kHead ? head_count(r) : tail_count(r);
  int64_t total = 0;
  for (int t0 = 0; t0 < T; t0 += 64) {
    if (!kHead && r.mapped && r.sam) {
      const int e_lo = t0 > 2 ? t0 - 2 : 0, e_hi = t0 + 62 < r.n_ev + 1 ? t0 + 62 : r.n_ev + 1;
      if (e_lo < e_hi && (e_lo < r.ev_from || e_hi > r.ev_from + kMdEvents)) {
        int n_ev;
        r.ev_from = e_lo;
        md_walk(t, r, rs, sq, len, lane, ev, r.ev_from, &n_ev);
      }
    }
    const int tk = t0 + lane;
    int l = 0;
    Tok d = empty_tok();
    if (tk < T) {
      d = kHead ? head_tok(t, r, tk) : tail_tok(r, nm, ev, tk);
      l = tok_len(d);
    }
    const int off = wave_excl_scan(l, lane);
    if (kWrite && tk < T) tok_write(t, d, out + at + total + off);
    total += __shfl(off + l, 63);
  }
  return total;
}

// NAME, or QUAL (reversed when rev), at dst: the lanes' register bytes,
// then the rest from memory
__device__ __forceinline__ void put_field(uint32_t rb, const char *src, int n, bool rev, int lane, char *dst) {
#pragma unroll
  for (int i = 0; i < kRegBytes; ++i) {
    const int k = lane + 64 * i;
    if (k < n) dst[rev ? n - 1 - k : k] = (char)reg_byte(rb, i);
  }
  for (int k = lane + 64 * kRegBytes; k < n; k += 64) dst[rev ? n - 1 - k : k] = src[k];
}

// All records of read q, at out (kWrite); returns their length.
template <bool kWrite>
__device__ __forceinline__ int64_t format_read(const TextArgs &t, int q, int lane, MdEvent *ev, char *out) {
  const int64_t n0 = t.name_off[q], q0 = t.qual_off[q];
  const int nl = (int)(t.name_off[q + 1] - n0), ql = (int)(t.qual_off[q + 1] - q0);
  const char *name = t.names + n0;
  const char *qual = t.quals + q0;
  const uint8_t *sq = t.seq + t.off[q];
  const int len = t.off[q + 1] - t.off[q];
  const int64_t a0 = t.aln_off[q], a1 = t.aln_off[q + 1];
  // the read's SEQ (both passes: the MD walk reads it), NAME and QUAL, in
  // registers, all loads issued together
  uint32_t rs = 0, rn = 0, rq = 0;
#pragma unroll
  for (int i = 0; i < kRegBytes; ++i) {
    const int k = lane + 64 * i;
    if (k < len) rs |= (uint32_t)sq[k] << (8 * i);
    if (kWrite && k < nl) rn |= (uint32_t)(uint8_t)name[k] << (8 * i);
    if (kWrite && k < ql) rq |= (uint32_t)(uint8_t)qual[k] << (8 * i);
  }
  const bool mapped = a1 > a0;
  const int n_rec = mapped ? (int)(a1 - a0) : 1;
  int64_t at = 0;
  for (int j = 0; j < n_rec; ++j) {
    Rec r{};
    r.mapped = mapped;
    r.sam = t.mode == 1;
    int nm = 0;
    if (mapped) {
      r.a = t.alns[a0 + j];
      if (r.sam) {
        r.rev = r.a.rb >= t.l_pac;
        sam_place(t, r.a, r.rev, r.fpos, r.beg, r.endc);
        r.clip_l = r.rev ? len - r.a.qe : r.a.qb;
        r.clip_r = r.rev ? r.a.qb : len - r.a.qe;
        r.flag = (r.rev ? 0x10 : 0) | (j > 0 ? 0x800 : 0);
        nm = md_walk(t, r, rs, sq, len, lane, ev, 0, &r.n_ev);
      }
    }
    if (kWrite) put_field(rn, name, nl, false, lane, out + at);
    at += nl;
    at += emit_tokens<kWrite, true>(t, r, rs, sq, len, nm, ev, lane, out, at);
    if (kWrite) {
      // SEQ (reverse-complemented for the reverse strand), the separator, QUAL
      char *s = out + at;
#pragma unroll
      for (int i = 0; i < kRegBytes; ++i) {
        const int k = lane + 64 * i;
        const int c = reg_byte(rs, i);
        if (k < len) s[r.rev ? len - 1 - k : k] = base_char(r.rev && c < 4 ? 3 - c : c);
      }
      for (int k = lane + 64 * kRegBytes; k < len; k += 64) {
        const int c = sq[k];
        s[r.rev ? len - 1 - k : k] = base_char(r.rev && c < 4 ? 3 - c : c);
      }
      if (lane == 0) s[len] = r.sam ? '\t' : ' ';
      put_field(rq, qual, ql, r.rev, lane, s + len + 1);
    }
    at += len + 1 + ql;
    at += emit_tokens<kWrite, false>(t, r, rs, sq, len, nm, ev, lane, out, at);
    __builtin_amdgcn_wave_barrier();  // the next record's events replace these
  }
  return at;
}

// The counting pass (len_out[q]: read q's text length) or the writing pass
// (read q's text at txt + txt_off[q]).  At least 4 waves per SIMD: the
// writing pass fits 128 registers without spills.
template <bool kWrite>
__global__ void __launch_bounds__(64 * kFmtWaves) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_fmt(TextArgs t, int64_t *len_out, const int64_t *txt_off, char *txt) {
  __shared__ MdEvent ev_all[kFmtWaves][kMdEvents];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the read's index, held in scalar registers (the same in every lane), so
  // that what follows from it is loaded once per wavefront
  const int q = __builtin_amdgcn_readfirstlane(blockIdx.x * kFmtWaves + wv);
  if (q >= t.n_reads) return;  // the whole wavefront
  const int64_t n = format_read<kWrite>(t, q, lane, ev_all[wv], kWrite ? txt + txt_off[q] : nullptr);
  if (len_out && lane == 0) len_out[q] = n;
}

// An upper bound on the text of each read, from what is known before the
// records are formatted.  A record is its NAME, SEQ and QUAL, the contig's
// name, at most 12 bytes a CIGAR operation (a count of at most 28 bits and
// the letter; the SAM core operations are a subset of n_core), the MD string
// and 256 bytes for every other field (numbers of at most 20 characters,
// tags and separators).  The MD string is at most 2 span + 1 bytes: an event
// writes the number of matches before it (at most one digit more than
// there are matches) and one base or '^' and the deleted bases, so never
// more than twice the reference bases it covers, plus one for the last run.
__global__ void __launch_bounds__(256) k_text_bound(TextArgs t, int64_t *bound) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= t.n_reads) return;
  const int64_t base = (t.name_off[q + 1] - t.name_off[q]) + (t.qual_off[q + 1] - t.qual_off[q]) +
                       (t.off[q + 1] - t.off[q]) + 256;
  const int64_t a0 = t.aln_off[q], a1 = t.aln_off[q + 1];
  int64_t u = a1 > a0 ? 0 : base;
  for (int64_t j = a0; j < a1; ++j) {
    const Aln &a = t.alns[j];
    u += base + (t.ctg_name_off[a.rid + 1] - t.ctg_name_off[a.rid]) + 12 * ((int64_t)a.n_cigar + a.n_core + 4) +
         2 * (a.re - a.rb) + 1;
  }
  bound[q] = (u + 3) & ~(int64_t)3;  // staging offsets stay 4-byte aligned
}

// flag |= some read's text is longer than its bound
__global__ void __launch_bounds__(256) k_text_check(int n, const int64_t *len, const int64_t *bound, int64_t *flag) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n && len[q] > bound[q]) atomicOr((unsigned long long *)flag, 1ull);
}

// read q's text from the staging buffer (4-byte aligned) to its place, 16
// lanes a read (a wavefront moves four reads at once: the copy is a few
// words a lane, so its time is the loads' latency): each lane writes aligned
// words of the destination, each
// made of two source words shifted into place; the words at either end that
// the read shares with its neighbours are written byte by byte.  A source
// word read past the text stays inside the staging buffer (bounds are
// rounded up to 4 bytes and the buffer has 16 to spare).
__global__ void __launch_bounds__(256) k_text_pack(int n, const char *stage, const int64_t *stage_off,
                                                   const int64_t *len, const int64_t *txt_off, char *txt) {
  const int q = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (q >= n) return;
  const int lane = threadIdx.x & 15;
  const char *src = stage + stage_off[q];
  const int64_t d0 = txt_off[q], l = len[q], d1 = d0 + l;
  if (l <= 0) return;
  const uint32_t *sw = reinterpret_cast<const uint32_t *>(src);
  uint32_t *dw = reinterpret_cast<uint32_t *>(txt);
  const int64_t w0 = d0 >> 2, w1 = (d1 + 3) >> 2;
#pragma unroll 4
  for (int64_t k = w0 + lane; k < w1; k += 16) {
    const int64_t a = 4 * k;
    if (a >= d0 && a + 4 <= d1) {
      const int64_t o = a - d0;  // source offset of the word's first byte, >= 0
      const uint32_t lo = sw[o >> 2], hi = sw[(o >> 2) + 1];
      dw[k] = __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(o & 3));
    } else {
      for (int64_t x = a > d0 ? a : d0, e = a + 4 < d1 ? a + 4 : d1; x < e; ++x) txt[x] = src[x - d0];
    }
  }
}

// MAPQ and MD-string length of every alignment (thread per alignment;
// aln_read gives the alignment's read)
__global__ void __launch_bounds__(256) k_aln_mapq_md(TextArgs t, int64_t n_alns, const int32_t *aln_read,
                                                     int32_t *mapq, int64_t *md_len) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alns) return;
  const Aln &a = t.alns[i];
  const int q = aln_read[i];
  const bool rev = a.rb >= t.l_pac;
  int64_t fpos;
  int beg, endc;
  sam_place(t, a, rev, fpos, beg, endc);
  CountSink c;
  emit_md(c, t, a, rev, t.seq + t.off[q], beg, endc, fpos);
  md_len[i] = c.n + 1;  // with its terminating zero
  mapq[i] = mapq_se(t, a.score, a.qb, a.qe, a.rb, a.re);
}

__global__ void __launch_bounds__(256) k_aln_md_write(TextArgs t, int64_t n_alns, const int32_t *aln_read,
                                                      const int64_t *md_off, char *md) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alns) return;
  const Aln &a = t.alns[i];
  const int q = aln_read[i];
  const bool rev = a.rb >= t.l_pac;
  int64_t fpos;
  int beg, endc;
  sam_place(t, a, rev, fpos, beg, endc);
  WriteSink w{md + md_off[i]};
  emit_md(w, t, a, rev, t.seq + t.off[q], beg, endc, fpos);
  w.put('\0');
}

}  // namespace

void launch_text(const TextArgs &t, int64_t *len, const int64_t *txt_off, char *txt, bool write, hipStream_t s) {
  if (t.n_reads <= 0) return;
  const unsigned blocks = (unsigned)((t.n_reads + kFmtWaves - 1) / kFmtWaves);
  if (write)
    hipLaunchKernelGGL(k_fmt<true>, dim3(blocks), dim3(64 * kFmtWaves), 0, s, t, len, txt_off, txt);
  else
    hipLaunchKernelGGL(k_fmt<false>, dim3(blocks), dim3(64 * kFmtWaves), 0, s, t, len, txt_off, txt);
}

void launch_text_staged(const TextArgs &t, int64_t *len, const int64_t *stage_off, char *stage, hipStream_t s) {
  if (t.n_reads <= 0) return;
  const unsigned blocks = (unsigned)((t.n_reads + kFmtWaves - 1) / kFmtWaves);
  hipLaunchKernelGGL(k_fmt<true>, dim3(blocks), dim3(64 * kFmtWaves), 0, s, t, len, stage_off, stage);
}
void launch_text_bound(const TextArgs &t, int64_t *bound, hipStream_t s) {
  if (t.n_reads > 0)
    hipLaunchKernelGGL(k_text_bound, dim3((unsigned)((t.n_reads + 255) / 256)), dim3(256), 0, s, t, bound);
}
void launch_text_check(int n, const int64_t *len, const int64_t *bound, int64_t *flag, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_text_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, len, bound, flag);
}
void launch_text_pack(int n, const char *stage, const int64_t *stage_off, const int64_t *len, const int64_t *txt_off,
                      char *txt, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_text_pack, dim3((unsigned)((n + 15) / 16)), dim3(256), 0, s, n, stage, stage_off, len,
                       txt_off, txt);
}

void launch_aln_mapq_md(const TextArgs &t, int64_t n_alns, const int32_t *aln_read, int32_t *mapq, int64_t *md_len,
                        hipStream_t s) {
  if (n_alns > 0)
    hipLaunchKernelGGL(k_aln_mapq_md, dim3((unsigned)((n_alns + 255) / 256)), dim3(256), 0, s, t, n_alns, aln_read,
                       mapq, md_len);
}

void launch_aln_md_write(const TextArgs &t, int64_t n_alns, const int32_t *aln_read, const int64_t *md_off,
                         char *md, hipStream_t s) {
  if (n_alns > 0)
    hipLaunchKernelGGL(k_aln_md_write, dim3((unsigned)((n_alns + 255) / 256)), dim3(256), 0, s, t, n_alns, aln_read,
                       md_off, md);
}

}  // namespace g3a
