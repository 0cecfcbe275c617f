// Per-read device logic of the BWA-MEM hot path for MI355X (gfx950).
//
// Everything here is __host__ __device__ so that the same code can also be
// compiled by g++ for CPU-side debugging; the product always runs it inside
// the HIP kernels of g3a_kernels.hip.  Each function cites the reference
// file:line whose behaviour it reproduces (paths relative to the reference).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define G3A_HD __host__ __device__ __forceinline__
#if defined(__HIP_DEVICE_COMPILE__)
#define G3A_POPC(x) __popcll(x)
#else
#define G3A_POPC(x) __builtin_popcountll(x)
#endif
#else
#define G3A_HD inline
#define G3A_POPC(x) __builtin_popcountll(x)
#endif

namespace g3a {

constexpr int kMaxReadLen = 320;  // MAX_LEN_READ, bwa/src/macro.h:27

// Options used on the path (subset of mem_opt_t, bwa/src/bwa.h:157-189).
struct Opt {
  int a, b, o_del, e_del, o_ins, e_ins, w, T, pen_clip5, pen_clip3;
  int min_seed_len, split_width, max_occ, max_chain_gap, flag;
  int64_t max_mem_intv;
  float split_factor, mask_level, drop_ratio, mask_level_redun;
  int8_t mat[25];
};

// Learned-index form: a two-stage model from the 32-symbol prefix of a
// suffix (2 bits per symbol, first symbol highest; symbols past the text end
// are 0) to its row.  The root is a piecewise-linear interpolation of the
// exact row of every 6-symbol prefix; it routes a key to one of kLiLeaves
// leaves.  A leaf is a one-hidden-layer network of the key with FP32
// weights: ReLU units relu(key - K_h) with their knots K_h at keys of the
// leaf, output y_0 + s_0 (key - K_0) + sum_h (s_h - s_{h-1}) relu(key - K_h)
// = the predicted row offset from the leaf's first row, a continuous
// piecewise-linear function of the key.  It is stored as its knots (K_k, the
// output y_k there and the slope s_k after it) and evaluated on the segment
// that holds the key as y_k + s_k (key - K_k), the difference taken in
// 64-bit integers before it is converted to FP32: the same function, with
// the resolution of each segment instead of the leaf's (a cluster of
// near-equal keys gets a segment of its own) and without adding up terms of
// opposite sign.  The knots of all leaves share one pool.
constexpr int kLiRootBits = 12;
constexpr int kLiLeaves = 4096;
constexpr int kLiWindowCap = 256;  // largest half-width of a search window
constexpr int kLiShort = 32;       // rows whose suffix has fewer than 32 symbols ('$' row included)

struct alignas(16) LiLeaf {
  uint32_t a;    // first row routed to the leaf
  uint32_t err;  // largest |predicted - true row| over the leaf's distinct keys
  uint32_t k0;   // first knot in the pool
  uint32_t nk;   // knots (hidden units + 1); 0 for a leaf without rows
};

struct alignas(16) LiKnot {
  uint64_t key;  // K_k
  float y, s;    // output at K_k, slope (rows per key unit) after it
};

// Device view of the index.  `occ` holds one 32-byte line per 64 BWT rows:
// four uint32 base counts of the rows before the line, relative to its
// 2^32-row superblock (`super`), then the 64 rows' 2-bit BWT symbols in two
// uint64 words, first row in the highest bits; the '$' row is stored as 0
// and discounted.  One Occ lookup is two 16-byte loads of one line.
struct DevIndex {
  const uint32_t *occ;
  int64_t super[4][4];
  int64_t count[5];  // count[c] = 1 + #bases < c (the '$' row first), count[4] = n+1
  int64_t primary;   // row whose suffix starts at text position 0 (BWT '$')
  int64_t n;         // text length 2*l_pac; rows 0..n
  int64_t l_pac;
  const uint8_t *pac;       // forward strand, 2 bits per base, MSB first
  const uint32_t *sa32;     // dense SA sample (one of sa32/sa64 is set)
  const uint64_t *sa64;
  int sa_shift;             // sample interval = 1 << sa_shift
  int n_contigs;
  const int64_t *ctg_off;   // n_contigs
  const int64_t *ctg_len;
  const int32_t *ctg_alt;
  // bidirectional interval (k, l, s) of every kKmer-mer, 2-bit code of the
  // first base highest; null when the text has 2^32 rows or more
  const uint32_t *kmer;
  // texts of fewer than 2^32 rows: 16-byte lines of 32 rows (21-bit counts
  // of bases 0-2 relative to a 2^21-row superblock, then 32 symbols) and the
  // superblocks' four absolute counts (uint32 each); null otherwise
  const uint32_t *occn;
  const uint32_t *sbc;
  // learned-index form (all null otherwise; texts of fewer than 2^32 rows):
  // the full suffix array, the 32-symbol key of every row, the root's
  // 2^kLiRootBits + 1 segment rows, the leaves and their knot pool, and the
  // (row, suffix length) pairs of the kLiShort short suffixes in row order
  const uint32_t *saf;
  const uint64_t *keys;
  const uint32_t *li_root;
  const LiLeaf *li_leaf;
  const LiKnot *li_knot;
  const uint32_t *li_short;
  uint32_t li_leaf_w;  // routing: leaf = predicted row / li_leaf_w
};

constexpr int kKmer = 12;

// Bidirectional interval: rows [k, k+s) of the text, rows [l, l+s) for the
// reverse complement; read[m..n].  P is int64_t for any text and uint32_t
// for texts of fewer than 2^32 rows (the seeding kernel's narrow form).
template <class P>
struct IntvT {
  P k, l, s;
  int32_t m, n;
};
using Intv = IntvT<int64_t>;
using IntvN = IntvT<uint32_t>;

template <class Q, class P>
G3A_HD IntvT<Q> intv_cast(const IntvT<P> &a) {
  IntvT<Q> r;
  r.k = (Q)a.k;
  r.l = (Q)a.l;
  r.s = (Q)a.s;
  r.m = a.m;
  r.n = a.n;
  return r;
}

// v[c] for c in 0..3 as a chain of selects: with v in scalar registers (a
// kernel argument) a per-lane c would otherwise turn into a memory load
template <class T>
G3A_HD T pick4(const T *v, int c) {
  return c == 0 ? v[0] : c == 1 ? v[1] : c == 2 ? v[2] : v[3];
}

struct Seed {
  int64_t rbeg;
  int32_t qbeg, len;
  int32_t rid, pad;
};

// ---------------------------------------------------------------------------
// FM-index primitives (bwa/src/fmindex.cu:29-45, 107-180; macro.h:50-71)
// ---------------------------------------------------------------------------
// Occurrences of each base in rows [line start, line start + y] from the two
// symbol words: bit planes hi/lo of the even bit of each pair.
G3A_HD void line_counts(uint64_t w0, uint64_t w1, int y, uint32_t o[4]) {
  const uint64_t E = 0x5555555555555555ULL;
  uint64_t m0, m1;
  if (y < 32) {
    m0 = (~0ULL << (62 - 2 * y)) & E;
    m1 = 0;
  } else {
    m0 = E;
    m1 = (~0ULL << (62 - 2 * (y - 32))) & E;
  }
  const uint64_t h0 = (w0 >> 1) & m0, l0 = w0 & m0, h1 = (w1 >> 1) & m1, l1 = w1 & m1;
  const uint32_t n3 = (uint32_t)(G3A_POPC(h0 & l0) + G3A_POPC(h1 & l1));
  const uint32_t n2 = (uint32_t)(G3A_POPC(h0 & ~l0) + G3A_POPC(h1 & ~l1));
  const uint32_t n1 = (uint32_t)(G3A_POPC(l0 & ~h0) + G3A_POPC(l1 & ~h1));
  o[3] = n3;
  o[2] = n2;
  o[1] = n1;
  o[0] = (uint32_t)(y + 1) - n1 - n2 - n3;
}

// The BWA .bwt body (bwt.c:385-419 layout): per 128 positions of the BWT
// string ('$' row left out), four uint64 counts of the positions before the
// block (eight uint32 words), then eight words of 16 symbols, the first
// symbol in the highest bits.  Row r is string position r below the '$' row
// `primary` and r - 1 above it.
// Counts of each base at string positions [0, k).
G3A_HD void bwa_occ_before(const uint32_t *W, uint64_t k, uint64_t o[4]) {
  const uint32_t *blk = W + (k >> 7) * 16;
#pragma unroll
  for (int x = 0; x < 4; ++x) o[x] = (uint64_t)blk[2 * x] | (uint64_t)blk[2 * x + 1] << 32;
  uint32_t rem = (uint32_t)(k & 127);
  for (int j = 0; rem > 0; ++j) {
    const uint32_t w = blk[8 + j], t = rem < 16 ? rem : 16;
    const uint32_t m = t == 16 ? 0x55555555u : 0x55555555u & ~(0xffffffffu >> (2 * t));
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const uint32_t d = w ^ (0x55555555u * (uint32_t)x);  // symbols equal to x become 00
      o[x] += (uint64_t)G3A_POPC((uint64_t)(~(d | d >> 1) & m));
    }
    rem -= t;
  }
}
// Base of string position k.
G3A_HD uint32_t bwa_sym(const uint32_t *W, uint64_t k) {
  return (W[(k >> 7) * 16 + 8 + ((k & 127) >> 4)] >> ((15 - (k & 15)) << 1)) & 3u;
}
// Counts of each base in rows [0, r) ('$' not counted).
G3A_HD void bwa_occ_rows(const uint32_t *W, uint64_t primary, uint64_t r, uint64_t o[4]) {
  bwa_occ_before(W, r <= primary ? r : r - 1, o);
}

constexpr int kNarrowSbShift = 21;  // narrow lines count relative to 2^21-row superblocks

// Line L of the Occ tables from the BWA .bwt body W (rows 0..n).  Lines
// 0 .. rows/64 are the wide lines (64 rows, 32 B: four uint32 counts
// relative to the line's 2^32-row superblock, then the symbols as two
// uint64 words, first row highest); when `narrow` is set, lines
// rows/64 + 1 .. are the narrow lines (32 rows, 16 B: counts of bases 0-2
// relative to the 2^21-row superblock at bits 0-20, 21-41, 42-62, then the
// symbols), and the first narrow line of each superblock writes its four
// absolute counts to sbc.  The '$' row is stored as 0 and never counted.
G3A_HD void occ_line_from_bwa(const uint32_t *W, uint64_t n, uint64_t primary, uint64_t L, uint32_t *wide,
                              uint32_t *narrow, uint32_t *sbc) {
  const uint64_t rows = n + 1, n_wide = (rows >> 6) + 1;
  const bool is_wide = L < n_wide;
  const uint64_t line = is_wide ? L : L - n_wide, span = is_wide ? 64 : 32;
  const uint64_t r0 = line * span;
  const uint64_t sb0 = is_wide ? r0 & ~0xffffffffULL : r0 & ~((1ULL << kNarrowSbShift) - 1);
  uint64_t c[4], b[4];
  bwa_occ_rows(W, primary, r0, c);
  bwa_occ_rows(W, primary, sb0, b);
  uint64_t w[2] = {0, 0};
  for (uint64_t y = 0; y < span; ++y) {
    const uint64_t r = r0 + y;
    if (r >= rows) break;
    if (r == primary) continue;
    const uint64_t ch = bwa_sym(W, r < primary ? r : r - 1);
    w[y >> 5] |= ch << (62 - 2 * (y & 31));
  }
  if (is_wide) {
    uint32_t *o = wide + line * 8;
    for (int x = 0; x < 4; ++x) o[x] = (uint32_t)(c[x] - b[x]);
    o[4] = (uint32_t)w[0];
    o[5] = (uint32_t)(w[0] >> 32);
    o[6] = (uint32_t)w[1];
    o[7] = (uint32_t)(w[1] >> 32);
  } else {
    const uint64_t packed = (c[0] - b[0]) | (c[1] - b[1]) << 21 | (c[2] - b[2]) << 42;
    uint32_t *o = narrow + line * 4;
    o[0] = (uint32_t)packed;
    o[1] = (uint32_t)(packed >> 32);
    o[2] = (uint32_t)w[0];
    o[3] = (uint32_t)(w[0] >> 32);
    if (r0 == sb0) {
      uint32_t *q = sbc + (r0 >> kNarrowSbShift) * 4;
      for (int x = 0; x < 4; ++x) q[x] = (uint32_t)c[x];
    }
  }
}

// One Occ line: the four counts and the two symbol words.
struct OccLine {
  uint32_t c[4];
  uint64_t w0, w1;
};
G3A_HD OccLine occ_line(const DevIndex &x, int64_t line) {
  const uint32_t *b = x.occ + (line << 3);
  OccLine L;
#pragma unroll
  for (int k = 0; k < 4; ++k) L.c[k] = b[k];
  L.w0 = (uint64_t)b[4] | (uint64_t)b[5] << 32;
  L.w1 = (uint64_t)b[6] | (uint64_t)b[7] << 32;
  return L;
}
// Occ(p, c) for all c from p's line
G3A_HD void occ_at(const DevIndex &x, const OccLine &L, int64_t p, int64_t o[4]) {
  uint32_t in[4];
  line_counts(L.w0, L.w1, (int)(p & 63), in);
  const int64_t *sb = x.super[p >> 32];
  const int dollar = (p >> 6) == (x.primary >> 6) && p >= x.primary;
#pragma unroll
  for (int c = 0; c < 4; ++c) o[c] = sb[c] + (int64_t)L.c[c] + in[c];
  o[0] -= dollar;
}

// Narrow form (texts of fewer than 2^32 rows, 32-bit arithmetic): one
// 16-byte line of 32 rows, plus the line's superblock counts.
struct OccLineN {
  uint64_t cnt, w;
};
G3A_HD OccLineN occn_line(const DevIndex &x, uint32_t line) {
  OccLineN L;
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4 v = *(const uint4 *)(x.occn + ((uint64_t)line << 2));
  L.cnt = (uint64_t)v.x | (uint64_t)v.y << 32;
  L.w = (uint64_t)v.z | (uint64_t)v.w << 32;
#else
  const uint32_t *b = x.occn + ((uint64_t)line << 2);
  L.cnt = (uint64_t)b[0] | (uint64_t)b[1] << 32;
  L.w = (uint64_t)b[2] | (uint64_t)b[3] << 32;
#endif
  return L;
}
// Occ(p, c) for all c from p's narrow line
G3A_HD void occ_at(const DevIndex &x, const OccLineN &L, uint32_t p, uint32_t o[4]) {
  const int y = (int)(p & 31);
  const uint64_t m = (~0ULL << (62 - 2 * y)) & 0x5555555555555555ULL;
  const uint64_t hi = (L.w >> 1) & m, lo = L.w & m;
  const uint32_t n3 = (uint32_t)G3A_POPC(hi & lo), n2 = (uint32_t)G3A_POPC(hi & ~lo),
                 n1 = (uint32_t)G3A_POPC(lo & ~hi);
  uint32_t s0, s1, s2, s3;
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4 sb = *(const uint4 *)(x.sbc + ((p >> 21) << 2));
  s0 = sb.x;
  s1 = sb.y;
  s2 = sb.z;
  s3 = sb.w;
#else
  const uint32_t *sb = x.sbc + ((p >> 21) << 2);
  s0 = sb[0];
  s1 = sb[1];
  s2 = sb[2];
  s3 = sb[3];
#endif
  const uint32_t r0 = (uint32_t)(L.cnt & 0x1fffff), r1 = (uint32_t)(L.cnt >> 21) & 0x1fffff,
                 r2 = (uint32_t)(L.cnt >> 42) & 0x1fffff;
  const uint32_t pr = (uint32_t)x.primary, ls = p & ~31u, ss = p & ~((1u << 21) - 1);
  // the '$' row is stored as symbol 0 and counted by no base
  const uint32_t dol_before = pr >= ss && pr < ls, dol_in = pr >= ls && pr <= p;
  const uint32_t r3 = (ls - ss) - r0 - r1 - r2 - dol_before;
  o[0] = s0 + r0 + ((uint32_t)y + 1 - n1 - n2 - n3) - dol_in;
  o[1] = s1 + r1 + n1;
  o[2] = s2 + r2 + n2;
  o[3] = s3 + r3 + n3;
}

// Occ(p, c): occurrences of c in rows [0, p], '$' excluded
G3A_HD void occ4(const DevIndex &x, int64_t p, int64_t o[4]) { occ_at(x, occ_line(x, p >> 6), p, o); }

// Occ at two rows; one line fetch when both lie in the same line (the
// common case once an interval has narrowed to a few rows)
G3A_HD void occ4_pair(const DevIndex &x, int64_t p1, int64_t p2, int64_t o1[4], int64_t o2[4]) {
  OccLine L = occ_line(x, p1 >> 6);
  occ_at(x, L, p1, o1);
  if ((p2 >> 6) != (p1 >> 6)) L = occ_line(x, p2 >> 6);
  occ_at(x, L, p2, o2);
}
// Narrow form: both lines are requested before either is used, so that when
// they differ the two fetches overlap (the same line twice is a cache hit)
G3A_HD void occ4_pair(const DevIndex &x, uint32_t p1, uint32_t p2, uint32_t o1[4], uint32_t o2[4]) {
  const OccLineN L1 = occn_line(x, p1 >> 5), L2 = occn_line(x, p2 >> 5);
  occ_at(x, L1, p1, o1);
  occ_at(x, L2, p2, o2);
}

G3A_HD int64_t occ1(const DevIndex &x, int64_t p, int c) {
  int64_t o[4];
  occ4(x, p, o);
  return o[c];
}

// BWT character of row k (4 for the '$' row)
G3A_HD int bwt_char(const DevIndex &x, int64_t k) {
  if (k == x.primary) return 4;
  const uint32_t *b = x.occ + ((k >> 6) << 3);
  const int y = (int)(k & 63);
  // rows 0-15 are the high half of the first uint64 word, i.e. uint32 b[5]
  const uint32_t w = b[4 + ((y >> 4) ^ 1)];
  return (int)((w >> (30 - 2 * (y & 15))) & 3);
}

template <class P>
G3A_HD void intv_init(const DevIndex &x, IntvT<P> &p, int c) {
  const P k = (P)pick4(x.count, c), l = (P)pick4(x.count, 3 - c), e = (P)pick4(x.count + 1, c);
  p.k = k;
  p.l = l;
  p.s = e - k;
}

// backwardExt (fmindex.cu:107-160): extension by `base` with the l update
template <class P>
G3A_HD void backward_ext(const DevIndex &x, const IntvT<P> &in, int base, IntvT<P> &out) {
  P osp[4], oep[4];
  occ4_pair(x, (P)(in.k - 1), (P)(in.k + in.s - 1), osp, oep);
  P s[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) s[b] = oep[b] - osp[b];
  const P pr = (P)x.primary;
  const P so = (in.k <= pr && in.k + in.s > pr) ? 1 : 0;
  const P l3 = in.l + so, l2 = l3 + s[3], l1 = l2 + s[2], l0 = l1 + s[1];
  const P l = base == 0 ? l0 : base == 1 ? l1 : base == 2 ? l2 : l3;
  out.k = (P)pick4(x.count, base) + pick4(osp, base);
  out.l = l;
  out.s = pick4(s, base);
}

// Occ(p, c) of one base c from p's narrow line: one popcount of the rows
// [line start, p] holding c ('$' is stored as 0 and not counted)
G3A_HD uint32_t occ1_at(const DevIndex &x, const OccLineN &L, uint32_t p, int c) {
  const int y = (int)(p & 31);
  const uint64_t m = (~0ULL << (62 - 2 * y)) & 0x5555555555555555ULL;
  const uint64_t hi = (L.w >> 1) & m, lo = L.w & m;
  const uint64_t sel = c == 3 ? hi & lo : c == 2 ? hi & ~lo : c == 1 ? lo & ~hi : m & ~hi & ~lo;
  const uint32_t n = (uint32_t)G3A_POPC(sel);
  uint32_t sc;
#if defined(__HIP_DEVICE_COMPILE__)
  const uint4 sb = *(const uint4 *)(x.sbc + ((p >> 21) << 2));
  sc = c == 0 ? sb.x : c == 1 ? sb.y : c == 2 ? sb.z : sb.w;
#else
  sc = x.sbc[((p >> 21) << 2) + c];
#endif
  const uint32_t r0 = (uint32_t)(L.cnt & 0x1fffff), r1 = (uint32_t)(L.cnt >> 21) & 0x1fffff,
                 r2 = (uint32_t)(L.cnt >> 42) & 0x1fffff;
  const uint32_t pr = (uint32_t)x.primary, ls = p & ~31u, ss = p & ~((1u << 21) - 1);
  const uint32_t dol_before = pr >= ss && pr < ls, dol_in = pr >= ls && pr <= p;
  const uint32_t rc = c == 0 ? r0 : c == 1 ? r1 : c == 2 ? r2 : (ls - ss) - r0 - r1 - r2 - dol_before;
  return sc + rc + n - (c == 0 ? dol_in : 0u);
}

// Backward extension by `base` of k and s only (l is left as it was): for
// the backward phases, whose intervals are only extended backward again or
// output, and neither reads l.  Narrow layout: one base's count per row.
G3A_HD void backward_ext_ks(const DevIndex &x, const IntvT<uint32_t> &in, int base, IntvT<uint32_t> &out) {
  const uint32_t p1 = in.k - 1, p2 = in.k + in.s - 1;
  const OccLineN L1 = occn_line(x, p1 >> 5), L2 = occn_line(x, p2 >> 5);  // overlapping fetches
  const uint32_t a = occ1_at(x, L1, p1, base);
  const uint32_t b = occ1_at(x, L2, p2, base);
  out.k = (uint32_t)pick4(x.count, base) + a;
  out.l = in.l;
  out.s = b - a;
}
G3A_HD void backward_ext_ks(const DevIndex &x, const IntvT<int64_t> &in, int base, IntvT<int64_t> &out) {
  backward_ext(x, in, base, out);
}

// backwardExtBackward (fmindex.cu:162-180): k and s only
G3A_HD void backward_ext_b(const DevIndex &x, Intv &io, int base) {
  int64_t oa[4], ob[4];
  occ4_pair(x, io.k - 1, io.k + io.s - 1, oa, ob);
  io.k = pick4(x.count, base) + pick4(oa, base);
  io.s = pick4(ob, base) - pick4(oa, base);
}

// FORWARD_EXT (macro.h:64-71)
template <class P>
G3A_HD void forward_ext(const DevIndex &x, IntvT<P> &io, int base) {
  IntvT<P> sw = io, o;
  sw.k = io.l;
  sw.l = io.k;
  backward_ext(x, sw, 3 - base, o);
  io.k = o.l;
  io.l = o.k;
  io.s = o.s;
}

// SA value of row k through the dense sample (LF walk to a sampled row)
G3A_HD int64_t sa_value(const DevIndex &x, int64_t k, int64_t *steps) {
  const int64_t mask = ((int64_t)1 << x.sa_shift) - 1;
  int64_t t = 0;
  while (k & mask) {
    if (k == x.primary) return t;
    int c = bwt_char(x, k);
    k = x.count[c] - 1 + occ1(x, k, c);
    ++t;
  }
  if (steps) *steps += t;
  int64_t i = k >> x.sa_shift;
  return t + (x.sa32 ? (int64_t)x.sa32[i] : (int64_t)x.sa64[i]);
}

// The same on the narrow layout (texts of fewer than 2^32 rows): each LF
// step reads the row's symbol and its Occ from one 16-byte line.
G3A_HD int64_t sa_value_n(const DevIndex &x, uint32_t k) {
  const uint32_t mask = (1u << x.sa_shift) - 1, pr = (uint32_t)x.primary;
  int64_t t = 0;
  while (k & mask) {
    if (k == pr) return t;
    const OccLineN L = occn_line(x, k >> 5);
    const int c = (int)((L.w >> (62 - 2 * (k & 31))) & 3);
    uint32_t o[4];
    occ_at(x, L, k, o);
    k = (uint32_t)pick4(x.count, c) - 1 + pick4(o, c);
    ++t;
  }
  const uint32_t i = k >> x.sa_shift;
  return t + (x.sa32 ? (int64_t)x.sa32[i] : (int64_t)x.sa64[i]);
}

// ---------------------------------------------------------------------------
// Learned-index form: text words, the model, exact windowed search
// ---------------------------------------------------------------------------
G3A_HD uint32_t bswap32_(uint32_t v) { return __builtin_bswap32(v); }

// Forward-strand symbols [i, i + 32), first in the highest bits.  pac is
// MSB-first and 4-byte aligned, with at least 16 bytes of padding.
G3A_HD uint64_t fwd_word(const DevIndex &x, int64_t i) {
  const uint64_t byte = (uint64_t)i >> 2;
  const uint32_t *w = (const uint32_t *)(x.pac + (byte & ~(uint64_t)3));
  const uint64_t be01 = (uint64_t)bswap32_(w[0]) << 32 | bswap32_(w[1]);
  const uint32_t be2 = bswap32_(w[2]);
  const int sh = (int)(byte & 3) * 8 + (int)(i & 3) * 2;  // < 32
  return sh ? (be01 << sh) | ((uint64_t)be2 >> (32 - sh)) : be01;
}

// The 32 two-bit symbols of v in reverse order
G3A_HD uint64_t rev_sym(uint64_t v) {
  v = ((v >> 2) & 0x3333333333333333ULL) | ((v & 0x3333333333333333ULL) << 2);
  v = ((v >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((v & 0x0F0F0F0F0F0F0F0FULL) << 4);
  return __builtin_bswap64(v);
}

// Text symbols [q, q + 32) for l_pac <= q < n (reverse-complement half:
// T[q + t] = 3 - fwd[n - 1 - q - t]); 0 past the text end.
G3A_HD uint64_t rc_word(const DevIndex &x, int64_t q) {
  const int64_t m = x.n - q;  // symbols left, 1..l_pac
  const uint64_t f = m >= 32 ? fwd_word(x, m - 32) : fwd_word(x, 0) >> (2 * (32 - m));
  uint64_t r = ~rev_sym(f);
  if (m < 32) r &= ~0ULL << (2 * (32 - m));
  return r;
}

// Text symbols [p, p + 32) of T = forward + reverse complement; 0 past the
// end.  This is the key of the row whose suffix starts at p.
G3A_HD uint64_t text_word(const DevIndex &x, int64_t p) {
  if (p >= x.n) return 0;
  if (p >= x.l_pac) return rc_word(x, p);
  const int64_t f = x.l_pac - p;
  const uint64_t w = fwd_word(x, p);
  if (f >= 32) return w;
  return (w & (~0ULL << (2 * (32 - f)))) | rc_word(x, x.l_pac) >> (2 * f);
}

// Root: the leaf of a key.  Rows of the 6-symbol prefix segment s are
// [li_root[s], li_root[s + 1]); the key's place in it is the next 32 key
// bits, interpolated in integer arithmetic (routing is exact and identical
// wherever it runs).  Monotone in the key, so each leaf holds a row range.
G3A_HD uint32_t li_route(const DevIndex &x, uint64_t key) {
  const uint32_t sg = (uint32_t)(key >> (64 - kLiRootBits));
  const uint32_t b0 = x.li_root[sg], b1 = x.li_root[sg + 1];
  const uint64_t frac = (key << kLiRootBits) >> 32;
  const uint64_t pred = b0 + (((uint64_t)(b1 - b0) * frac) >> 32);
  const uint64_t j = pred / x.li_leaf_w;
  return j < (uint64_t)kLiLeaves ? (uint32_t)j : (uint32_t)(kLiLeaves - 1);
}

// The leaf network's output: predicted row offset from the leaf's first row
// (the segment of the last knot <= key, found by bisection; keys below the
// first knot extend its segment)
G3A_HD float li_leaf_eval(const LiKnot *pool, const LiLeaf &L, uint64_t key) {
  if (L.nk == 0) return 0.f;
  const LiKnot *kn = pool + L.k0;
  int lo = 0, hi = (int)L.nk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (kn[mid].key <= key) lo = mid;
    else hi = mid - 1;
  }
  const LiKnot k = kn[lo];
  const float d = key >= k.key ? (float)(key - k.key) : -(float)(k.key - key);
  return k.y + k.s * d;
}

G3A_HD int64_t li_predict(const LiKnot *pool, const LiLeaf &L, uint64_t key) {
  float y = li_leaf_eval(pool, L, key);
  y = y > -1e9f ? (y < 5e9f ? y : 5e9f) : -1e9f;
  return (int64_t)L.a + (int64_t)(y >= 0.f ? y + 0.5f : y - 0.5f);
}

// First row whose key is >= K, searched exactly in the window of half-width
// min(err + 2, kLiWindowCap) around the model's prediction; -1 when the
// window does not bracket it (the caller then takes the FM-index path).
G3A_HD int64_t li_lower_bound(const DevIndex &x, uint64_t K) {
  const LiLeaf &L = x.li_leaf[li_route(x, K)];
  const int64_t N = x.n + 1;
  const int64_t p = li_predict(x.li_knot, L, K);
  int64_t E = (int64_t)L.err + 2;
  if (E > kLiWindowCap) E = kLiWindowCap;
  int64_t lo = p - E, hi = p + E + 1;
  if (lo < 0) lo = 0;
  if (hi > N) hi = N;
  if (lo >= hi) return -1;
  if (lo > 0 && x.keys[lo - 1] >= K) return -1;
  if (hi < N && x.keys[hi] < K) return -1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (x.keys[mid] < K) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Exact search in [lo, hi): first row whose key is >= K, or hi
G3A_HD int64_t li_search(const DevIndex &x, uint64_t K, int64_t lo, int64_t hi) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (x.keys[mid] < K) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Buckets of more rows than this are narrowed by the model's window
constexpr int64_t kLiModelMin = 64;

// Rows [k0, k1) hold the keys that start with the j-symbol string W.
// Suffixes shorter than j whose zero-padded key starts with W do not start
// with W; they sort first in [k0, k1) ('$' is below every base).  Their
// padding makes W's last symbol A, so only then is their table read.
G3A_HD bool li_finish(const DevIndex &x, uint64_t W, int j, int64_t k0, int64_t k1, uint32_t &k, uint32_t &s) {
  int64_t kk = k0;
  if (((W >> (64 - 2 * j)) & 3) == 0 && k1 > k0) {
    for (int t = 0; t < kLiShort; ++t) {
      const int64_t r = x.li_short[2 * t];
      if (r >= k0 && r < k1 && (int)x.li_short[2 * t + 1] < j) ++kk;
    }
  }
  k = (uint32_t)kk;
  s = (uint32_t)(k1 - kk);
  return true;
}

// Rows [k, k + s) of the suffixes that start with the j-symbol string W
// (first symbol in the highest bits, the rest of the word 0; 1 <= j <= 32):
// the interval the FM-index's backward search gives for it.  For j >= kKmer
// the k-mer table's interval of W's first kKmer symbols bounds the search
// exactly; in a bucket of more than kLiModelMin rows the model's window
// (checked against the keys at its edges) narrows it further.  Without the
// table the model's window alone is searched, and false is returned when it
// misses.  lc (optional): [4] lookups the model narrowed, [5] model windows
// that missed.
G3A_HD bool li_range(const DevIndex &x, uint64_t W, int j, uint32_t &k, uint32_t &s, int64_t *lc = nullptr) {
  const int64_t N = x.n + 1;
  const uint64_t unit = (uint64_t)1 << (64 - 2 * j);
  const uint64_t Wn = W + unit;  // first key past the prefix; 0 when W is all T
  int64_t lo = 0, hi = N;        // lb(W) lies in [lo, hi]; rows >= hi do not start with W
  int64_t bend = N;
  bool bounded = false;
  if (x.kmer && j >= kKmer) {
    const uint32_t *e = x.kmer + 3 * (uint64_t)(W >> (64 - 2 * kKmer));
    lo = e[0];
    hi = lo + e[2];
    bend = hi;
    bounded = true;
    if (lo == hi) {  // the first kKmer symbols do not occur
      k = (uint32_t)lo;
      s = 0;
      return true;
    }
  }
  if (!bounded || hi - lo > kLiModelMin) {
    const LiLeaf &L = x.li_leaf[li_route(x, W)];
    const int64_t p = li_predict(x.li_knot, L, W);
    int64_t E = (int64_t)L.err + 2;
    if (E > kLiWindowCap) E = kLiWindowCap;
    int64_t wl = p - E, wh = p + E + 1;
    wl = wl < lo ? lo : wl;
    wh = wh > hi ? hi : wh;
    const bool ok = wl < wh && (wl == lo || x.keys[wl - 1] < W) && (wh == hi || x.keys[wh] >= W);
    if (ok) {
      lo = wl;
      hi = wh;
      if (lc) ++lc[4];
    } else {
      if (lc) ++lc[5];
      if (!bounded) return false;
    }
  }
  const int64_t k0 = li_search(x, W, lo, hi);
  // end of the run: gallop from k0 over rows k0 + 2^i - 1, then bisect
  int64_t k1;
  const int64_t end = bend;
  if (Wn == 0) {
    k1 = end;
  } else {
    int64_t prev = k0 - 1, r = k0, step = 1;
    while (r < end && x.keys[r] < Wn && step < 1024) {
      prev = r;
      step <<= 1;
      r = k0 + step - 1;
    }
    if (r < end && x.keys[r] < Wn) {
      if (bounded) {
        k1 = li_search(x, Wn, r + 1, end);
      } else {
        k1 = li_lower_bound(x, Wn);
        if (k1 < 0) return false;
      }
    } else {
      k1 = li_search(x, Wn, prev + 1, r < end ? r : end);
    }
  }
  return li_finish(x, W, j, k0, k1, k, s);
}

// Full-SA construction step: from sampled row j << sa_shift, whose value
// is known, walk back through the rows of positions p - 1, p - 2, ... by LF
// steps, writing each, until the next sampled row or position 0 (the row
// whose BWT symbol is '$').  Every row is written by exactly one walk.
// False if the walk leaves the text (an inconsistent index).
G3A_HD bool li_walk_sample(const DevIndex &x, int64_t j, uint32_t *saf) {
  const int64_t r64 = j << x.sa_shift;
  if (r64 > x.n) return true;
  uint32_t r = (uint32_t)r64;
  int64_t p = x.sa32 ? (int64_t)x.sa32[j] : (int64_t)x.sa64[j];
  saf[r] = (uint32_t)p;
  const uint32_t mask = (1u << x.sa_shift) - 1, pr = (uint32_t)x.primary;
  while (r != pr) {
    const OccLineN L = occn_line(x, r >> 5);
    const int c = (int)((L.w >> (62 - 2 * (r & 31))) & 3);
    uint32_t o[4];
    occ_at(x, L, r, o);
    r = (uint32_t)pick4(x.count, c) - 1 + pick4(o, c);
    --p;
    if ((r & mask) == 0) break;
    if (p < 0 || r > (uint32_t)x.n) return false;
    saf[r] = (uint32_t)p;
  }
  return true;
}

// Number of leading positions u < maxlen with q[u] a base equal to the text
// symbol T[t + u], t + u < n: how far a match known to occur once, ending
// just before t, continues forward.
G3A_HD int text_run(const DevIndex &x, int64_t t, const uint8_t *q, int maxlen) {
  const int64_t lim = x.n - t;
  if (lim < maxlen) maxlen = (int)lim;
  int u = 0;
  while (u < maxlen) {
    const int c = maxlen - u < 32 ? maxlen - u : 32;
    uint64_t rw = 0;
    int v = 0;
    for (; v < c; ++v) {
      const int b = q[u + v];
      if (b > 3) break;
      rw |= (uint64_t)b << (62 - 2 * v);
    }
    const uint64_t tw = text_word(x, t + u);
    const uint64_t d = v ? (tw ^ rw) & (~0ULL << (64 - 2 * v)) : 0;
    if (d) return u + (__builtin_clzll(d) >> 1);
    u += v;
    if (v < c) return u;
  }
  return u;
}

// bns_pos2rid_gpu / bns_intv2rid_gpu (bwa/src/bntseq.cu:7-37)
G3A_HD int pos2rid(const DevIndex &x, int64_t pos_f) {
  if (pos_f >= x.l_pac) return -1;
  int left = 0, mid = 0, right = x.n_contigs;
  while (left < right) {
    mid = (left + right) >> 1;
    if (pos_f >= x.ctg_off[mid]) {
      if (mid == x.n_contigs - 1) break;
      if (pos_f < x.ctg_off[mid + 1]) break;
      left = mid + 1;
    } else right = mid;
  }
  return mid;
}
G3A_HD int64_t depos(const DevIndex &x, int64_t pos, int &is_rev) {
  is_rev = pos >= x.l_pac;
  return is_rev ? (x.l_pac << 1) - 1 - pos : pos;
}
G3A_HD int intv2rid(const DevIndex &x, int64_t rb, int64_t re) {
  int is_rev;
  if (rb < x.l_pac && re > x.l_pac) return -2;
  int rid_b = pos2rid(x, depos(x, rb, is_rev));
  int rid_e = rb < re ? pos2rid(x, depos(x, re - 1, is_rev)) : rid_b;
  return rid_b == rid_e ? rid_b : -1;
}

// base of the fwd+revcomp text at position p
G3A_HD int text_base_p(const uint8_t *pac, int64_t l_pac, int64_t p) {
  if (p < l_pac) return (pac[p >> 2] >> ((~p & 3) << 1)) & 3;
  int64_t q = (l_pac << 1) - 1 - p;
  return 3 - ((pac[q >> 2] >> ((~q & 3) << 1)) & 3);
}
G3A_HD int text_base(const DevIndex &x, int64_t p) { return text_base_p(x.pac, x.l_pac, p); }

// Sequential reader of the fwd+revcomp text: keeps one 32-bit word (16
// bases) of the packed forward strand, so a run of neighbouring positions
// costs one load per 16 bases.  The packed strand must be readable up to the
// next 4-byte boundary.
struct TextReader {
  const uint8_t *pac;
  int64_t l_pac;
  int64_t widx;
  uint32_t word;
  G3A_HD TextReader(const uint8_t *p, int64_t l) : pac(p), l_pac(l), widx(-1), word(0) {}
  G3A_HD int operator()(int64_t p) {
    const bool fwd = p < l_pac;
    const int64_t q = fwd ? p : (l_pac << 1) - 1 - p;
    const int64_t wi = q >> 4;
    if (wi != widx) {
      widx = wi;
      const uint8_t *b = pac + (wi << 2);
#if defined(__HIP_DEVICE_COMPILE__)
      word = *(const uint32_t *)b;
#else
      word = (uint32_t)b[0] | (uint32_t)b[1] << 8 | (uint32_t)b[2] << 16 | (uint32_t)b[3] << 24;
#endif
    }
    const int v = (int)((word >> (8 * ((q >> 2) & 3) + 2 * (3 - (q & 3)))) & 3);
    return fwd ? v : 3 - v;
  }
};

// Sequential reader of a byte array (read codes): keeps the 16-byte aligned
// chunk holding the last byte read, so a run of neighbouring positions costs
// one vector load per 16 bytes.  The array must be readable up to the next
// 16-byte boundary past its end (the device read buffer has that padding).
struct ByteReader {
  const uint8_t *base;
  int64_t cidx;
  uint32_t w0, w1, w2, w3;
  G3A_HD explicit ByteReader(const uint8_t *b) : base(b), cidx(-1), w0(0), w1(0), w2(0), w3(0) {}
  G3A_HD int operator()(int j) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint8_t *p = base + j;
    const uint64_t a = (uint64_t)p;
    const int64_t ci = (int64_t)(a >> 4);
    if (ci != cidx) {
      cidx = ci;
      const uint4 v = *(const uint4 *)(p - (a & 15));
      w0 = v.x;
      w1 = v.y;
      w2 = v.z;
      w3 = v.w;
    }
    const int k = (int)((a >> 2) & 3);
    const uint32_t x = k == 0 ? w0 : k == 1 ? w1 : k == 2 ? w2 : w3;
    return (int)((x >> (8 * (a & 3))) & 0xff);
#else
    return base[j];
#endif
  }
};

// Scores of one target base against query codes 0..4, packed: codes 0..3 as
// bytes of lo, code 4 in hi.  Replaces per-cell lookups into the score
// matrix (a kernel argument, so a per-lane index would be a memory load).
struct ScoreRow {
  uint32_t lo;
  int hi;
  G3A_HD int operator()(int q) const { return q < 4 ? (int)(int8_t)(uint8_t)(lo >> (8 * q)) : hi; }
};
G3A_HD ScoreRow score_row(const Opt &o, int t) {
  ScoreRow r;
  r.lo = 0;
  for (int c = 0; c < 4; ++c) {
    const uint32_t v0 = (uint8_t)o.mat[c], v1 = (uint8_t)o.mat[5 + c], v2 = (uint8_t)o.mat[10 + c],
                   v3 = (uint8_t)o.mat[15 + c], v4 = (uint8_t)o.mat[20 + c];
    const uint32_t v = t == 0 ? v0 : t == 1 ? v1 : t == 2 ? v2 : t == 3 ? v3 : v4;
    r.lo |= v << (8 * c);
  }
  r.hi = t == 0 ? o.mat[4] : t == 1 ? o.mat[9] : t == 2 ? o.mat[14] : t == 3 ? o.mat[19] : o.mat[24];
  return r;
}

// ---------------------------------------------------------------------------
// Seeding
// ---------------------------------------------------------------------------

// Output sink for one read's intervals: the first `cap` go to `slot`, the
// rest are only counted (the read is then redone with a bigger buffer).
struct IntvSink {
  Intv *slot;
  int cap, n;
  int64_t n_seed;  // sum over intervals of min(s, max_occ)
  int max_occ;
  template <class P>
  G3A_HD void push(const IntvT<P> &p) {
    if (n < cap) slot[n] = intv_cast<int64_t>(p);
    ++n;
    n_seed += (int64_t)p.s > max_occ ? max_occ : (int64_t)p.s;
  }
};

// Storage of the LEP and found lists.  LepWide keeps the whole interval;
// LepNarrow packs it into 16 bytes (one vector load or store per entry) and
// holds for texts of fewer than 2^32 rows.
// Each names the interval type its searches compute in (I).
struct LepWide {
  using I = Intv;
  static constexpr bool kNarrow = false;
  Intv v;
  static G3A_HD Intv get(const LepWide *p, int i) { return p[i].v; }
  static G3A_HD void put(LepWide *p, int i, const Intv &x) { p[i].v = x; }
};
struct alignas(16) LepNarrow {
  using I = IntvN;
  static constexpr bool kNarrow = true;
  uint32_t k, l, s;
  int16_t m, n;
  static G3A_HD IntvN get(const LepNarrow *p, int i) {
    const LepNarrow e = p[i];
    IntvN r;
    r.k = e.k;
    r.l = e.l;
    r.s = e.s;
    r.m = e.m;
    r.n = e.n;
    return r;
  }
  static G3A_HD void put(LepNarrow *p, int i, const IntvN &x) {
    LepNarrow e;
    e.k = x.k;
    e.l = x.l;
    e.s = x.s;
    e.m = (int16_t)x.m;
    e.n = (int16_t)x.n;
    p[i] = e;
  }
};

// Scratch for the LEP list and the found list.
// A list in a global slice of entries of type E.
template <class E>
struct GlobalList {
  using I = typename E::I;
  E *p;
  G3A_HD I get(int i) const { return E::get(p, i); }
  G3A_HD void put(int i, const I &v) const { E::put(p, i, v); }
};

// A narrow list whose first `capl` entries live in LDS and the rest in a
// global slice.  In LDS an entry is three words, k, s and m | n << 16, the
// words of entry i at (3i + c) * stride from the lane's first word (lanes
// interleaved).  l is not kept: these lists only hold intervals that are
// extended backward or output, and neither reads l.
struct SplitList {
  using I = IntvN;
  uint32_t *lds;
  int stride, capl;
  LepNarrow *g;
  G3A_HD I get(int i) const {
    if (i >= capl) return LepNarrow::get(g, i - capl);
    const uint32_t *w = lds + 3 * i * stride;
    I r;
    r.k = w[0];
    r.l = 0;
    r.s = w[stride];
    const uint32_t mn = w[2 * stride];
    r.m = (int32_t)(mn & 0xffff);
    r.n = (int32_t)(mn >> 16);
    return r;
  }
  G3A_HD void put(int i, const I &v) const {
    if (i >= capl) return LepNarrow::put(g, i - capl, v);
    uint32_t *w = lds + 3 * i * stride;
    w[0] = v.k;
    w[stride] = v.s;
    w[2 * stride] = (uint32_t)(v.m & 0xffff) | (uint32_t)v.n << 16;
  }
};

// The LEP list a and the found list f of one thread, with their capacities.
template <class LA, class LF>
struct LepBuf {
  using I = typename LA::I;
  LA a;
  LF f;
  int cap, cap_f;
};
template <class E>
using GlobalBuf = LepBuf<GlobalList<E>, GlobalList<E>>;

// The backward steps of rounds 1 and 2, one list entry after another.  A
// step rebuilds its list in place (entry nc is written only after every
// entry j >= nc has been read) and returns the new length, or -1 when a list
// outgrows its capacity.  Other policies (a wavefront's lanes over the
// entries) give the same lists.
struct SerialBack {
  // Round 1 at read position i (base c, -1 past the read or at an N): an
  // entry that extends goes on, deduplicated by size; the first that does
  // not, when none before it did, is a new start, stored in `fl` when the
  // match is long enough to be emitted.
  template <class L, class LF>
  static G3A_HD int round1(const DevIndex &x, const Opt &o, const L &list, int np, int i, int c, const LF &fl,
                           int &nf, int cap, int cap_f, int &last_start, int64_t *steps) {
    using I = typename L::I;
    using P = decltype(I::s);
    int nc = 0;
    P last_s = 0;  // size of the last entry written (read only when nc > 0)
    for (int j = 0; j < np; ++j) {
      const I pj = list.get(j);
      I ok;
      bool ext_ok = false;
      if (c >= 0) {
        backward_ext_ks(x, pj, c, ok);
        if (steps) ++*steps;
        ext_ok = ok.s >= 1;
      }
      if (!ext_ok) {
        if (nc == 0 && i + 1 < last_start) {
          last_start = i + 1;
          if (pj.n - i >= o.min_seed_len) {
            I p = pj;
            p.m = i + 1;
            if (nf == cap_f) return -1;
            fl.put(nf++, p);
          }
        }
      } else if (nc == 0 || ok.s != last_s) {
        ok.m = pj.m;
        ok.n = pj.n;
        if (nc == cap) return -1;
        list.put(nc++, ok);
        last_s = ok.s;
      }
    }
    return nc;
  }
  // Round 2 at pivot-side position m (base = read[m - 1]): an extension at
  // least as large as the last kept one goes on; the first entry that falls
  // short, when long enough, becomes a seed.
  template <class L, class LF, class P>
  static G3A_HD int reseed(const DevIndex &x, const Opt &o, const L &leps, int num_old, int m, int base,
                           P min_seed_intv, const LF &seeds, int &ns, int cap_s, int64_t *steps) {
    using I = typename L::I;
    P curr = min_seed_intv;
    bool collected = false;
    int nl = 0;
    for (int j = 0; j < num_old; ++j) {
      I lep = leps.get(j), ext;
      backward_ext_ks(x, lep, base, ext);
      if (steps) ++*steps;
      if (ext.s >= curr) {
        ext.m = m - 1;
        ext.n = lep.n;
        leps.put(nl++, ext);
        curr = ext.s;
      } else if (!collected && lep.n - lep.m + 1 >= o.min_seed_len) {
        if (ns == cap_s) return -1;
        seeds.put(ns++, lep);
        collected = true;
      }
    }
    return nl;
  }
};

// Round 1 (preseedAndFilterV2, bwa/src/seed.cu:25-104 with
// preseedOnePos2Backward, seed.cu:454-531).  The reference keeps, for each
// end n, the longest match read[m(n)..n] and outputs it when it is right
// maximal; these are exactly the super-maximal exact matches of length >=
// min_seed_len.  They are found here with the bidirectional SMEM search
// (forward from a pivot collecting the points where the interval shrinks,
// then backward), which visits each read position a few times instead of
// once per end position.  Output order is by end position, as the
// reference's.  Returns false when a LEP list outgrows `buf`.
// kLearn = false compiles the learned-index form out (the FM-only kernel
// keeps its registers); with kLearn it is used when x.saf / x.li_leaf are set.
template <bool kLearn = true, class B, class BW = SerialBack>
G3A_HD bool smem_round1(const DevIndex &x, const Opt &o, const uint8_t *q, int len, IntvSink &out, B &buf,
                        int64_t *steps, int64_t *lc = nullptr) {
  using I = typename B::I;
  using P = decltype(I::s);
  int x0 = 0;
  while (x0 < len) {
    if (q[x0] > 3) { ++x0; continue; }
    // the backward pass rebuilds the list in place: entry nc is written only
    // after entry j >= nc has been read
    const auto &prev = buf.a;
    int np = 0;
    I ik;
    intv_init(x, ik, q[x0]);
    ik.m = x0;
    ik.n = x0;
    int i = x0 + 1;
    if (kLearn && x0 == 0 && x.kmer && len >= kKmer) {
      // the pass from 0 has no backward extensions: only its longest match
      // is used, so it may start from the k-mer table's interval of
      // read[0, kKmer) when that occurs (the shorter prefixes' entries
      // would never be read)
      uint32_t code = 0;
      int t = 0;
      for (; t < kKmer && q[t] < 4; ++t) code = code << 2 | q[t];
      if (t == kKmer) {
        const uint32_t *e = x.kmer + 3 * (uint64_t)code;
        if (e[2] > 0) {
          ik.k = (P)e[0];
          ik.l = (P)e[1];
          ik.s = (P)e[2];
          ik.n = kKmer - 1;
          i = kKmer;
        }
      }
    }
    for (; i < len; ++i) {
      if (kLearn && x.saf && ik.s == 1) {
        // one occurrence, at saf[k]: the forward extensions keep s = 1 for
        // as long as the read matches the text after it, so compare the two
        // directly (only l, which nothing downstream reads, is left stale)
        const int run = text_run(x, (int64_t)x.saf[ik.k] + (i - x0), q + i, len - i);
        if (lc) {
          ++lc[2];
          lc[3] += run;
        }
        ik.n = i + run - 1;
        i += run;
        if (i < len) {
          if (np == buf.cap) return false;
          prev.put(np++, ik);
        }
        break;
      }
      if (q[i] < 4) {
        I ok = ik;
        forward_ext(x, ok, q[i]);
        if (steps) ++*steps;
        if (ok.s != ik.s) {
          if (np == buf.cap) return false;
          prev.put(np++, ik);
          if (ok.s < 1) break;
        }
        ik = ok;
        ik.n = i;
      } else {
        if (np == buf.cap) return false;
        prev.put(np++, ik);
        break;
      }
    }
    if (i == len) {
      if (np == buf.cap) return false;
      prev.put(np++, ik);
    }
    // longest first
    for (int a = 0, b = np - 1; a < b; ++a, --b) {
      const I ta = prev.get(a), tb = prev.get(b);
      prev.put(a, tb);
      prev.put(b, ta);
    }
    const int ret = prev.get(0).n + 1;
    // backward search; SMEMs are found in decreasing start order
    int nf = 0, last_start = len + 1;  // start of the last match found (none yet)
    const auto &fl = buf.f;
    for (i = x0 - 1; i >= -1; --i) {
      const int c = i < 0 ? -1 : (q[i] < 4 ? q[i] : -1);
      const int nc = BW::round1(x, o, prev, np, i, c, fl, nf, buf.cap, buf.cap_f, last_start, steps);
      if (nc < 0) return false;
      if (nc == 0) break;
      np = nc;
    }
    // found in decreasing start (and end) order: emit them by end
    for (int j = nf - 1; j >= 0; --j) out.push(fl.get(j));
    x0 = ret;
  }
  return true;
}

// Round 1 by the reference's own per-end backward search; needs no scratch.
// Used for reads whose LEP lists do not fit the fast path's buffers.
G3A_HD void smem_round1_per_end(const DevIndex &x, const Opt &o, const uint8_t *q, int len, IntvSink &out,
                                int64_t *steps) {
  int prev_m = -1, prev_valid = 0;
  Intv prev;
  prev.k = prev.l = prev.s = 0;
  prev.m = prev.n = 0;
  for (int n = 0; n < len; ++n) {
    int valid = 0;
    Intv cur;
    cur.k = cur.l = cur.s = 0;
    cur.m = cur.n = 0;
    if (n >= o.min_seed_len - 1 && q[n] < 4) {
      intv_init(x, cur, q[n]);
      int m = n;
      while (m > 0 && q[m - 1] < 4) {
        Intv t = cur;
        backward_ext_b(x, t, q[m - 1]);
        if (steps) ++*steps;
        if (t.s < 1) break;
        cur.k = t.k;
        cur.s = t.s;
        --m;
      }
      cur.m = m;
      cur.n = n;
      valid = n - m + 1 >= o.min_seed_len;
    }
    // previous end is kept unless this end extends the same start
    if (prev_valid && !(valid && cur.m == prev_m)) out.push(prev);
    prev = cur;
    prev_m = cur.m;
    prev_valid = valid;
  }
  if (prev_valid) out.push(prev);
}

// Round 2, reseedV2 (bwa/src/seed.cu:639-913), for one SMEM.  `leps` and
// `seeds` are scratch lists of capacities cap and cap_s.  Returns false on
// overflow.
template <class LA, class LF, class BW = SerialBack>
G3A_HD bool reseed_one(const DevIndex &x, const Opt &o, const uint8_t *q, int len, const Intv &smem,
                       const LA &leps, const LF &seeds, int cap, int cap_s, IntvSink &out, int64_t *steps) {
  using I = typename LA::I;
  using P = decltype(I::s);
  const int split_len = (int)(o.min_seed_len * o.split_factor + .499);
  const int slen = smem.n - smem.m + 1;
  if (slen < split_len || smem.s > o.split_width) return true;
  const P min_seed_intv = (P)(smem.s + 1);
  const int pivot = (smem.m + smem.n + 1) >> 1;
  if (pivot > 77) return true;  // seed.cu:718
  if (q[pivot] >= 4) return true;
  int nl = 0, ns = 0;
  I seed, old;
  intv_init(x, seed, q[pivot]);
  seed.m = pivot;
  seed.n = pivot;
  int n;
  for (n = pivot; n < len - 1; ++n) {
    old = seed;
    if (q[n + 1] >= 4) break;
    forward_ext(x, seed, q[n + 1]);
    if (steps) ++*steps;
    seed.n = n + 1;
    if (seed.s < old.s) {
      if (nl == cap) return false;
      leps.put(nl++, old);
      if (seed.s < min_seed_intv) break;
    }
  }
  if (seed.s >= min_seed_intv) {
    seed.n = n;
    if (nl == cap) return false;
    leps.put(nl++, seed);
  }
  for (int a = 0, b = nl - 1; a < b; ++a, --b) {
    const I ta = leps.get(a), tb = leps.get(b);
    leps.put(a, tb);
    leps.put(b, ta);
  }
  for (int m = pivot; m > 0; --m) {
    const int base = q[m - 1];
    if (base >= 4) break;
    nl = BW::reseed(x, o, leps, nl, m, base, min_seed_intv, seeds, ns, cap_s, steps);
    if (nl < 0) return false;
    if (nl == 0) break;
  }
  if (nl > 0) {
    const I l0 = leps.get(0);
    if (l0.n - l0.m + 1 >= o.min_seed_len) {
      if (ns == cap_s) return false;
      seeds.put(ns++, l0);
    }
  }
  for (int j = 0; j < ns; ++j) out.push(seeds.get(j));
  return true;
}

// Round 3, reseedThirdRound (bwa/src/seed.cu:1143-1190) via reseedLastRound
// (seed.cu:1518-1550).  Every start m extends forward at least to length
// min_len before it can emit, so when the first kKmer bases hold no N their
// interval is read from the k-mer table instead of kKmer-1 forward steps.
template <class P, bool kLearn = true>
G3A_HD void reseed_round3(const DevIndex &x, const Opt &o, const uint8_t *q, int len, IntvSink &out,
                          int64_t *steps, int64_t *lc = nullptr) {
  const int min_len = o.min_seed_len + 1;
  const int64_t max_intv = o.max_mem_intv;
  const bool table = x.kmer && min_len > kKmer;
  const bool learned = kLearn && x.li_leaf && min_len <= 32;
  int m = 0;
  while (m < len) {
    int next_m = m + 1;
    int base = q[m];
    if (base < 4 && learned && m + min_len <= len) {
      // the pass emits at length min_len when fewer than max_intv rows hold
      // read[m, m + min_len): the learned index gives that interval directly
      uint64_t W = 0;
      int t = 0;
      for (; t < min_len && q[m + t] < 4; ++t) W |= (uint64_t)q[m + t] << (62 - 2 * t);
      uint32_t kk, ss;
      if (t == min_len) {
        if (lc) ++lc[0];
        if (li_range(x, W, min_len, kk, ss, lc)) {
          if ((int64_t)ss < max_intv) {
            if (ss > 0) {
              IntvT<P> sd;
              sd.k = (P)kk;
              sd.l = 0;
              sd.s = (P)ss;
              sd.m = m;
              sd.n = m + min_len - 1;
              out.push(sd);
            }
            m += min_len;
            continue;
          }
        } else if (lc) {
          ++lc[1];
        }
      }
    }
    if (base < 4) {
      IntvT<P> seed;
      int n0 = m + 1;
      uint32_t code = 0;
      bool jumped = false;
      if (table && m + kKmer <= len) {
        int k = 0;
        for (; k < kKmer && q[m + k] < 4; ++k) code = code << 2 | q[m + k];
        if (k == kKmer) {
          const uint32_t *e = x.kmer + 3 * (uint64_t)code;
          seed.k = (P)e[0];
          seed.l = (P)e[1];
          seed.s = (P)e[2];
          n0 = m + kKmer;
          next_m = n0;
          jumped = true;
        }
      }
      if (!jumped) intv_init(x, seed, base);
      for (int n = n0; n < len; ++n) {
        next_m = n + 1;
        base = q[n];
        if (base < 4) {
          forward_ext(x, seed, base);
          if (steps) ++*steps;
          if ((int64_t)seed.s < max_intv && (n - m + 1) >= min_len) {
            if (seed.s > 0) {
              seed.m = m;
              seed.n = n;
              out.push(seed);
            }
            break;
          }
        } else break;
      }
    }
    m = next_m;
  }
}

// Interval of k-mer `code` (kKmer bases, first base in the highest bits) by
// kKmer-1 forward steps from the first base: the k-mer table's entries.
G3A_HD void kmer_interval(const DevIndex &x, uint32_t code, uint32_t out[3]) {
  Intv seed;
  intv_init(x, seed, (int)(code >> (2 * (kKmer - 1))) & 3);
  for (int i = kKmer - 2; i >= 0; --i) forward_ext(x, seed, (int)(code >> (2 * i)) & 3);
  out[0] = (uint32_t)seed.k;
  out[1] = (uint32_t)seed.l;
  out[2] = (uint32_t)seed.s;
}

// ---------------------------------------------------------------------------
// Seed sort (sortSeedsLowDim/HighDim, bwa/src/aux.cu:1436-1610): stable by
// (qbeg, len) for <= 64 seeds; for more, the first 2048 positions are sorted
// stably by (qbeg, qbeg+len) and the rest stay in place.
// ---------------------------------------------------------------------------
G3A_HD bool seed_less(const Seed &a, const Seed &b) {
  return a.qbeg < b.qbeg || (a.qbeg == b.qbeg && a.len < b.len);
}
G3A_HD void sort_seeds(Seed *a, int n, Seed *tmp) {
  int m = n <= 64 ? n : (n < 2048 ? n : 2048);
  if (m <= 32) {
    for (int i = 1; i < m; ++i) {
      Seed v = a[i];
      int j = i - 1;
      while (j >= 0 && seed_less(v, a[j])) { a[j + 1] = a[j]; --j; }
      a[j + 1] = v;
    }
    return;
  }
  for (int width = 1; width < m; width *= 2) {
    for (int lo = 0; lo < m; lo += 2 * width) {
      int mid = lo + width < m ? lo + width : m, hi = lo + 2 * width < m ? lo + 2 * width : m;
      int i = lo, j = mid, k = lo;
      while (i < mid && j < hi) tmp[k++] = seed_less(a[j], a[i]) ? a[j++] : a[i++];
      while (i < mid) tmp[k++] = a[i++];
      while (j < hi) tmp[k++] = a[j++];
    }
    for (int i = 0; i < m; ++i) a[i] = tmp[i];
  }
}

// ---------------------------------------------------------------------------
// Chaining: BTreeChaining (bwa/src/aux.cu:1755-1825) over klib's B-tree
// (bwa/src/kbtree_CUDA.cu:3-144) with t = 6, the value kb_init_chn(512)
// yields for a 32-byte chain record.  Keys are chain ids; nodes live in a
// per-read arena.
// ---------------------------------------------------------------------------
constexpr int kKbT = 6, kKbN = 2 * kKbT - 1;
struct KbNode {
  int32_t n, internal;
  int32_t key[kKbN];
  int32_t child[kKbN + 1];
};
struct ChainTmp {  // chain under construction
  int64_t pos;
  int32_t rid, n, head, tail;  // seed list as links through `next`
};

struct KbTree {
  KbNode *nodes;
  int n_nodes, cap_nodes, root, n_keys;
  const ChainTmp *ch;
  G3A_HD int cmp(int64_t a, int64_t b) const { return (b < a) - (a < b); }
  G3A_HD int getp(const KbNode &xn, int64_t kpos, int *r) const {
    int tr, *rr = r ? r : &tr, begin = 0, end = xn.n;
    if (xn.n == 0) return -1;
    while (begin < end) {
      int mid = (begin + end) >> 1;
      if (cmp(ch[xn.key[mid]].pos, kpos) < 0) begin = mid + 1;
      else end = mid;
    }
    if (begin == xn.n) { *rr = 1; return xn.n - 1; }
    if ((*rr = cmp(kpos, ch[xn.key[begin]].pos)) < 0) --begin;
    return begin;
  }
  G3A_HD int lower(int64_t kpos) const {
    int low = -1, r = 0, xi = root;
    while (xi >= 0) {
      const KbNode &xn = nodes[xi];
      int i = getp(xn, kpos, &r);
      if (i >= 0 && r == 0) return xn.key[i];
      if (i >= 0) low = xn.key[i];
      if (!xn.internal) return low;
      xi = xn.child[i + 1];
    }
    return low;
  }
  G3A_HD int new_node(int internal) {
    if (n_nodes == cap_nodes) return -1;
    KbNode &z = nodes[n_nodes];
    z.n = 0;
    z.internal = internal;
    return n_nodes++;
  }
  G3A_HD bool split(int xi, int i, int yi) {
    int zi = new_node(nodes[yi].internal);
    if (zi < 0) return false;
    KbNode &xn = nodes[xi], &yn = nodes[yi], &zn = nodes[zi];
    zn.n = kKbT - 1;
    for (int a = 0; a < kKbT - 1; ++a) zn.key[a] = yn.key[kKbT + a];
    if (yn.internal)
      for (int a = 0; a < kKbT; ++a) zn.child[a] = yn.child[kKbT + a];
    yn.n = kKbT - 1;
    for (int a = xn.n; a >= i + 1; --a) xn.child[a + 1] = xn.child[a];
    xn.child[i + 1] = zi;
    for (int a = xn.n - 1; a >= i; --a) xn.key[a + 1] = xn.key[a];
    xn.key[i] = yn.key[kKbT - 1];
    ++xn.n;
    return true;
  }
  G3A_HD bool put(int cid) {
    ++n_keys;
    int64_t kpos = ch[cid].pos;
    int ri = root;
    if (nodes[ri].n == kKbN) {
      int si = new_node(1);
      if (si < 0) return false;
      nodes[si].child[0] = ri;
      root = si;
      if (!split(si, 0, ri)) return false;
      ri = si;
    }
    int xi = ri;
    for (;;) {
      KbNode &xn = nodes[xi];
      if (!xn.internal) {
        int i = getp(xn, kpos, nullptr);
        for (int a = xn.n - 1; a >= i + 1; --a) xn.key[a + 1] = xn.key[a];
        xn.key[i + 1] = cid;
        ++xn.n;
        return true;
      }
      int i = getp(xn, kpos, nullptr) + 1;
      if (nodes[xn.child[i]].n == kKbN) {
        if (!split(xi, i, xn.child[i])) return false;
        if (cmp(kpos, ch[nodes[xi].key[i]].pos) > 0) ++i;
      }
      xi = nodes[xi].child[i];
    }
  }
};


// test_and_merge (bwa/src/aux.cu:54-84); seeds of a chain are linked through nxt
G3A_HD int test_and_merge(const DevIndex &x, const Opt &o, ChainTmp &c, const Seed *s, int32_t *nxt, int pi) {
  const Seed &p = s[pi], &last = s[c.tail], &first = s[c.head];
  int64_t qend = last.qbeg + last.len, rend = last.rbeg + last.len;
  if (p.rid != c.rid) return 0;
  if (p.qbeg >= first.qbeg && p.qbeg + p.len <= qend && p.rbeg >= first.rbeg && p.rbeg + p.len <= rend) return 1;
  if ((last.rbeg < x.l_pac || first.rbeg < x.l_pac) && p.rbeg >= x.l_pac) return 0;
  int64_t xd = p.qbeg - last.qbeg, yd = p.rbeg - last.rbeg;
  if (yd >= 0 && xd - yd <= o.w && yd - xd <= o.w && xd - last.len < o.max_chain_gap &&
      yd - last.len < o.max_chain_gap) {
    nxt[c.tail] = pi;
    nxt[pi] = -1;
    c.tail = pi;
    ++c.n;
    return 1;
  }
  return 0;
}

// Builds the chains of one read from its sorted seeds.  Chain ids are
// written to `order` in B-tree order.  Returns the number of chains, or -1
// when the node arena is too small.
G3A_HD int chain_build(const DevIndex &x, const Opt &o, const Seed *s, int ns, ChainTmp *ch, int32_t *nxt,
                       KbNode *nodes, int cap_nodes, int32_t *order) {
  KbTree t;
  t.nodes = nodes;
  t.n_nodes = 0;
  t.cap_nodes = cap_nodes;
  t.n_keys = 0;
  t.ch = ch;
  t.root = t.new_node(0);
  if (t.root < 0) return -1;
  int nch = 0;
  for (int i = 0; i < ns; ++i) {
    const Seed &p = s[i];
    bool start_new = true;
    if (t.n_keys) {
      int low = t.lower(p.rbeg);
      if (low >= 0 && test_and_merge(x, o, ch[low], s, nxt, i)) start_new = false;
    }
    if (!start_new) continue;
    if (p.rid < 0 || p.rid >= 455) continue;  // bwa/src/aux.cu:1799
    ChainTmp &c = ch[nch];
    c.pos = p.rbeg;
    c.rid = p.rid;
    c.n = 1;
    c.head = c.tail = i;
    nxt[i] = -1;
    if (!t.put(nch)) return -1;
    ++nch;
  }
  // in-order traversal (__kb_traverse, kbtree_CUDA.cu:121-144); each frame
  // is (node, next key index, whether child `index` was already visited)
  int k = 0;
  if (nch && !nodes[t.root].internal) {
    // one leaf (most reads): its keys in order
    const KbNode &xn = nodes[t.root];
    for (int i = 0; i < xn.n; ++i) order[k++] = xn.key[i];
  } else if (nch) {
    int sn[40], si[40], sd[40], sp = 0;
    sn[0] = t.root; si[0] = 0; sd[0] = 0;
    while (sp >= 0) {
      const KbNode &xn = nodes[sn[sp]];
      const int i = si[sp];
      if (i > xn.n) { --sp; continue; }
      if (xn.internal && !sd[sp]) {
        sd[sp] = 1;
        ++sp;
        sn[sp] = xn.child[i]; si[sp] = 0; sd[sp] = 0;
        continue;
      }
      if (i < xn.n) order[k++] = xn.key[i];
      si[sp] = i + 1;
      sd[sp] = 0;
    }
  }
  return k;
}

// mem_chain_weight (bwa/src/aux.cu:91-110) over a linked chain
G3A_HD int chain_weight(const Seed *s, const int32_t *nxt, const ChainTmp &c) {
  int64_t end = 0;
  int w = 0, tmp;
  for (int j = c.head; j >= 0; j = nxt[j]) {
    const Seed &p = s[j];
    if (p.qbeg >= end) w += p.len;
    else if (p.qbeg + p.len > end) w += (int)(p.qbeg + p.len - end);
    end = end > p.qbeg + p.len ? end : p.qbeg + p.len;
  }
  tmp = w;
  w = 0;
  end = 0;
  for (int j = c.head; j >= 0; j = nxt[j]) {
    const Seed &p = s[j];
    if (p.rbeg >= end) w += p.len;
    else if (p.rbeg + p.len > end) w += (int)(p.rbeg + p.len - end);
    end = end > p.rbeg + p.len ? end : p.rbeg + p.len;
  }
  w = w < tmp ? w : tmp;
  return w < 1 << 30 ? w : (1 << 30) - 1;
}

// sortChainsDecreasingWeight + CHAINFILTERING_filter_kernel
// (bwa/src/aux.cu:1828-1988).  `order` holds chain ids in B-tree order and is
// rewritten with the kept chains in output order; `w16` and `tmp` are scratch
// of n entries.  Stable descending sort of the first 2048 chains when
// n <= 3072 (more chains stay unsorted); the filter is the sequential fixed
// point of the reference's spin-wait (i dropped iff an earlier kept chain j
// overlaps it by >= mask_level of the shorter and outweighs it); no filter
// above 4096 chains.  Returns the number of kept chains.
G3A_HD int chain_filter(const Opt &o, const Seed *s, const int32_t *nxt, const ChainTmp *ch, int32_t *order,
                        int n, uint16_t *w16, int32_t *tmp, uint8_t *kept) {
  if (n == 0) return 0;
  for (int i = 0; i < n; ++i) w16[i] = (uint16_t)chain_weight(s, nxt, ch[order[i]]);
  if (n <= 3072) {
    int m = n < 2048 ? n : 2048;
    // stable descending merge sort of (order, w16)[0..m)
    uint16_t *wt = (uint16_t *)(tmp + m);
    for (int width = 1; width < m; width *= 2) {
      for (int lo = 0; lo < m; lo += 2 * width) {
        int mid = lo + width < m ? lo + width : m, hi = lo + 2 * width < m ? lo + 2 * width : m;
        int i = lo, j = mid, k = lo;
        while (i < mid && j < hi) {
          if (w16[j] > w16[i]) { tmp[k] = order[j]; wt[k++] = w16[j++]; }
          else { tmp[k] = order[i]; wt[k++] = w16[i++]; }
        }
        while (i < mid) { tmp[k] = order[i]; wt[k++] = w16[i++]; }
        while (j < hi) { tmp[k] = order[j]; wt[k++] = w16[j++]; }
      }
      for (int i = 0; i < m; ++i) { order[i] = tmp[i]; w16[i] = wt[i]; }
    }
  }
  if (n > 4096) return n;
  for (int i = 0; i < n; ++i) {
    const ChainTmp &ci = ch[order[i]];
    uint16_t bi = (uint16_t)s[ci.head].qbeg, ei = (uint16_t)(s[ci.tail].qbeg + s[ci.tail].len);
    kept[i] = 1;
    for (int j = 0; j < i; ++j) {
      if (!kept[j]) continue;
      const ChainTmp &cj = ch[order[j]];
      uint16_t bj = (uint16_t)s[cj.head].qbeg, ej = (uint16_t)(s[cj.tail].qbeg + s[cj.tail].len);
      int b_max = bj > bi ? bj : bi, e_min = ej < ei ? ej : ei;
      if (e_min > b_max) {  // no ALT contigs are supported (is_alt == 0)
        int li = ei - bi, lj = ej - bj, min_l = li < lj ? li : lj;
        if (e_min - b_max >= min_l * o.mask_level && min_l < o.max_chain_gap) {
          if (w16[i] < w16[j] * o.drop_ratio && w16[j] - w16[i] >= o.min_seed_len << 1) {
            kept[i] = 0;
            break;
          }
        }
      }
    }
  }
  int k = 0;
  for (int i = 0; i < n; ++i)
    if (kept[i]) order[k++] = order[i];
  return k;
}

// cal_max_gap (bwa/src/aux.cu:168-175)
G3A_HD int cal_max_gap(const Opt &o, int qlen) {
  int l_del = (int)((double)(qlen * o.a - o.o_del) / o.e_del + 1.);
  int l_ins = (int)((double)(qlen * o.a - o.o_ins) / o.e_ins + 1.);
  int l = l_del > l_ins ? l_del : l_ins;
  l = l > 1 ? l : 1;
  return l < o.w << 1 ? l : o.w << 1;
}

// One SW extension task: a seed of a kept chain with the chain's reference
// window [rmax0, rmax1) (ExtendingPairGenerate, bwa/src/aux.cu:2101-2192).
struct Task {
  int64_t rbeg, rmax0, rmax1;
  int32_t qbeg, len, read, rid;
};

// Reference window of a chain, clipped to its strand and contig
// (aux.cu:2111-2132 with bns_fetch_seq_gpu, bwa/src/bntseq.cu:62-87).
G3A_HD void chain_window(const DevIndex &x, const Opt &o, const Seed *s, const int32_t *nxt, const ChainTmp &c,
                         int l_seq, int64_t &rmax0, int64_t &rmax1, int &rid) {
  const int64_t l_pac = x.l_pac;
  rmax0 = l_pac << 1;
  rmax1 = 0;
  for (int j = c.head; j >= 0; j = nxt[j]) {
    const Seed &p = s[j];
    int64_t b = p.rbeg - (p.qbeg + cal_max_gap(o, p.qbeg));
    int64_t e = p.rbeg + p.len + ((l_seq - p.qbeg - p.len) + cal_max_gap(o, l_seq - p.qbeg - p.len));
    rmax0 = rmax0 < b ? rmax0 : b;
    rmax1 = rmax1 > e ? rmax1 : e;
  }
  rmax0 = rmax0 > 0 ? rmax0 : 0;
  rmax1 = rmax1 < l_pac << 1 ? rmax1 : l_pac << 1;
  int64_t mid = s[c.head].rbeg;
  if (rmax0 < l_pac && l_pac < rmax1) {
    if (mid < l_pac) rmax1 = l_pac;
    else rmax0 = l_pac;
  }
  int is_rev;
  rid = pos2rid(x, depos(x, mid, is_rev));
  int64_t far_beg = x.ctg_off[rid], far_end = far_beg + x.ctg_len[rid];
  if (is_rev) {
    int64_t t = far_beg;
    far_beg = (l_pac << 1) - far_end;
    far_end = (l_pac << 1) - t;
  }
  rmax0 = rmax0 > far_beg ? rmax0 : far_beg;
  rmax1 = rmax1 < far_end ? rmax1 : far_end;
}

// Result of one extension task (mem_alnreg_t subset, bwa/src/bwa.h:229-248).
struct Region {
  int64_t rb, re;
  int32_t qb, qe, score, rid;
  int32_t seedlen0, secondary;
};

// filterRegions + sortRegions (bwa/src/aux.cu:2408-2613) for one read, in
// place; returns the number of kept regions.  The reference's threads race on
// the keep/secondary flags and write kept regions through an atomic counter;
// here anchors are taken in ascending index order and kept regions are
// written in ascending index order.  `flag`/`sec` are scratch of n entries.
G3A_HD int region_filter(const Opt &o, Region *a, int n, uint8_t *flag, int32_t *sec) {
  for (int i = 0; i < n; ++i) flag[i] = a[i].score >= o.T;
  for (int i = 0; i < n; ++i) {
    if (!flag[i]) continue;
    for (int j = i + 1; j < n; ++j) {
      if (!flag[j]) continue;
      if (a[i].rid != a[j].rid) continue;
      int64_t ai_rb = a[i].rb, ai_re = a[i].re, aj_rb = a[j].rb, aj_re = a[j].re;
      if (ai_rb < aj_rb && ai_re < aj_rb) continue;
      if (aj_rb < ai_rb && aj_re < ai_rb) continue;
      int or_ = (int)(ai_rb < aj_rb ? ai_re - aj_rb : aj_re - ai_rb);
      int oq = a[i].qb < a[j].qb ? a[i].qe - a[j].qb : a[j].qe - a[i].qb;
      int rli = (int)(ai_re - ai_rb), rlj = (int)(aj_re - aj_rb);
      int mr = rli < rlj ? rli : rlj;
      int mq = (a[i].qe - a[i].qb < a[j].qe - a[j].qb) ? a[i].qe - a[i].qb : a[j].qe - a[j].qb;
      if (or_ > o.mask_level_redun * mr && oq > o.mask_level_redun * mq) {
        if (a[i].score < a[j].score) { flag[i] = 0; break; }
        else flag[j] = 0;
      }
    }
  }
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (flag[i]) a[m++] = a[i];
  for (int i = 0; i < m; ++i) sec[i] = -1;
  for (int i = 0; i < m; ++i) {
    for (int j = i + 1; j < m; ++j) {
      if (sec[i] >= 0) break;
      if (sec[j] >= 0) continue;
      int qb_max = a[i].qb > a[j].qb ? a[i].qb : a[j].qb;
      int qe_min = a[i].qe < a[j].qe ? a[i].qe : a[j].qe;
      if (qe_min > qb_max) {
        int li = a[i].qe - a[i].qb, lj = a[j].qe - a[j].qb, min_l = li < lj ? li : lj;
        if (qe_min - qb_max >= min_l * o.mask_level) {
          if (a[i].score > a[j].score) sec[j] = i;
          else sec[i] = j;
        }
      }
    }
  }
  int k = 0;
  for (int i = 0; i < m; ++i) {
    if (sec[i] >= 0 && !(o.flag & 0x8)) continue;
    if (sec[i] >= 0 && a[i].score < a[sec[i]].score * o.drop_ratio) continue;
    Region r = a[i];
    r.secondary = sec[i];
    a[k++] = r;
  }
  // sortRegions: bubble sort by score descending (no ALT contigs)
  if (k == 2) {
    if (a[0].score < a[1].score) { Region t = a[0]; a[0] = a[1]; a[1] = t; }
  } else if (k > 2) {
    for (int i = 0; i < k - 1; ++i) {
      bool swapped = false;
      for (int j = 0; j < k - 1 - i; ++j)
        if (a[j].score < a[j + 1].score) { Region t = a[j]; a[j] = a[j + 1]; a[j + 1] = t; swapped = true; }
      if (!swapped) break;
    }
  }
  return k;
}

// infer_bw (bwa/src/aux.cu:506-513)
G3A_HD int infer_bw(int l1, int l2, int score, int a, int q, int r) {
  int w;
  if (l1 == l2 && l1 * a - score < (q + r - a) << 1) return 0;
  w = (int)((double)((l1 < l2 ? l1 : l2) * a - score - q) / r + 2.);
  int d = l1 > l2 ? l1 - l2 : l2 - l1;
  if (w < d) w = d;
  return w;
}

// Band width of the global alignment (FINALIZEALN_preprocessing2,
// bwa/src/aux.cu:2676-2752), capped at GLOBALSW_BANDWITH_CUTOFF = 500.
G3A_HD int global_bw(const Opt &o, int lq, int64_t rlen, int score) {
  if (lq == rlen) return 0;
  int w = infer_bw(lq, (int)rlen, score, o.a, o.o_del, o.e_del);
  int tmp = infer_bw(lq, (int)rlen, score, o.a, o.o_ins, o.e_ins);
  w = w > tmp ? w : tmp;
  int max_ins = (int)((double)(((lq + 1) >> 1) * o.a - o.o_ins) / o.e_ins + 1.);
  int max_del = (int)((double)(((lq + 1) >> 1) * o.a - o.o_del) / o.e_del + 1.);
  int max_gap = max_ins > max_del ? max_ins : max_del;
  max_gap = max_gap > 1 ? max_gap : 1;
  int d = (int)rlen > lq ? (int)rlen - lq : lq - (int)rlen;
  tmp = (max_gap + d + 1) >> 1;
  w = w < tmp ? w : tmp;
  tmp = d + 3;
  w = w > tmp ? w : tmp;
  return w < 500 ? w : 500;
}

G3A_HD int push_cigar(uint32_t *cigar, int n, int op, int len) {
  if (n == 0 || op != (int)(cigar[n - 1] & 0xf)) {
    cigar[n] = (uint32_t)(len << 4 | op);
    return n + 1;
  }
  cigar[n - 1] += (uint32_t)(len << 4);
  return n;
}

// ksw_global2 (bwa/src/ksw.cu:706-812): banded global alignment with
// traceback; the reference's ksw_global3 keeps int8_t DP state and a shared
// border buffer raced by 32 threads, so its defined intent (this routine) is
// used.  `tq`/`tt` give query/target bases; eh_h/eh_e hold qlen+1 entries and
// z holds tlen*min(qlen, 2w+1) direction bytes.  Returns the score.
// Row buffers of the banded global alignment: H and E of the previous row
// by query column.  EhArray keeps whole rows; EhRing keeps a ring of RB
// columns (enough when 2w+2 <= RB) at a stride, e.g. one lane's slice of a
// wavefront's LDS block.
struct EhArray {
  static constexpr bool kStage = false;
  int32_t *h, *e;
  G3A_HD int32_t &H(int j) { return h[j]; }
  G3A_HD int32_t &E(int j) { return e[j]; }
};
// The ring also carries `zst`, RB bytes (16-byte aligned) where a row's
// direction bytes are gathered before they are stored 16 bytes at a time:
// a thread's row is then whole 16-byte pieces of its alignment's matrix
// instead of 4-byte pieces, which the memory side merges far better when
// every lane of a wavefront writes its own alignment.
template <int RB>
struct EhRing {
  static constexpr bool kStage = true;
  int32_t *h, *e;
  int stride;
  uint32_t *zst;
  G3A_HD int32_t &H(int j) { return h[(j & (RB - 1)) * stride]; }
  G3A_HD int32_t &E(int j) { return e[(j & (RB - 1)) * stride]; }
};

// z-matrix row stride: the band's column count rounded up to 16 bytes
G3A_HD int z_stride(int qlen, int w) {
  const int n_col = qlen < 2 * w + 1 ? qlen : 2 * w + 1;
  return (n_col + 15) & ~15;
}

// ksw_global2 (bwa/src/ksw.cu:706-812, DR-9): banded global alignment.
// global_dp fills the direction bytes (a row's written four at a time, or
// gathered in eh.zst and written sixteen at a time) and returns the score;
// global_walk turns them into the CIGAR.
template <class Q, class T, class EH>
G3A_HD int global_dp(const Opt &o, int qlen, Q tq, int tlen, T tt, int w, EH eh, uint8_t *z) {
  const int MINUS_INF = -0x40000000;
  const int oe_del = o.o_del + o.e_del, oe_ins = o.o_ins + o.e_ins;
  const int zs = z_stride(qlen, w);
  eh.H(0) = 0;
  eh.E(0) = MINUS_INF;
  int j;
  for (j = 1; j <= qlen && j <= w; ++j) { eh.H(j) = -(o.o_ins + o.e_ins * j); eh.E(j) = MINUS_INF; }
  for (; j <= qlen && j <= w + 1; ++j) eh.H(j) = eh.E(j) = MINUS_INF;
  for (int i = 0; i < tlen; ++i) {
    int32_t f = MINUS_INF, h1, beg, end, t;
    const ScoreRow sr = score_row(o, tt(i));
    beg = i > w ? i - w : 0;
    end = i + w + 1 < qlen ? i + w + 1 : qlen;
    h1 = beg == 0 ? -(o.o_del + o.e_del * (i + 1)) : MINUS_INF;
    uint32_t *zi = (uint32_t *)&z[(int64_t)i * zs];
    uint32_t zw = 0;
    for (j = beg; j < end; ++j) {
      int32_t h, m = eh.H(j), e = eh.E(j);
      uint32_t d;
      eh.H(j) = h1;
      m += sr(tq(j));
      d = m >= e ? 0 : 1;
      h = m >= e ? m : e;
      d = h >= f ? d : 2;
      h = h >= f ? h : f;
      h1 = h;
      t = m - oe_del;
      e -= o.e_del;
      d |= e > t ? 1 << 2 : 0;
      e = e > t ? e : t;
      eh.E(j) = e;
      t = m - oe_ins;
      f -= o.e_ins;
      d |= f > t ? 2 << 4 : 0;
      f = f > t ? f : t;
      const int c = j - beg;
      zw |= d << (8 * (c & 3));
      if ((c & 3) == 3) {
        if constexpr (EH::kStage) eh.zst[c >> 2] = zw;
        else zi[c >> 2] = zw;
        zw = 0;
      }
    }
    if constexpr (EH::kStage) {
      if ((end - beg) & 3) eh.zst[(end - beg) >> 2] = zw;
      // the row's words, whole 16-byte pieces (bytes past the band are
      // never read)
#if defined(__HIPCC__)
      for (int q = 0; q < ((end - beg + 15) >> 4); ++q)
        reinterpret_cast<uint4 *>(zi)[q] = reinterpret_cast<const uint4 *>(eh.zst)[q];
#else
      for (int q = 0; q < ((end - beg + 15) >> 4) * 4; ++q) zi[q] = eh.zst[q];
#endif
    } else {
      if ((end - beg) & 3) zi[(end - beg) >> 2] = zw;
    }
    eh.H(end) = h1;
    eh.E(end) = MINUS_INF;
  }
  return eh.H(qlen);
}

// The traceback of global_dp's direction bytes: CIGAR (BAM encoding) into
// cigar, returns its operation count
G3A_HD int global_walk(int qlen, int tlen, int w, const uint8_t *z, uint32_t *cigar) {
  const int zs = z_stride(qlen, w);
  int n = 0, which = 0, i = tlen - 1, k = (i + w + 1 < qlen ? i + w + 1 : qlen) - 1;
  while (i >= 0 && k >= 0) {
    which = z[(int64_t)i * zs + (k - (i > w ? i - w : 0))] >> (which << 1) & 3;
    if (which == 0) { n = push_cigar(cigar, n, 0, 1); --i; --k; }
    else if (which == 1) { n = push_cigar(cigar, n, 2, 1); --i; }
    else { n = push_cigar(cigar, n, 1, 1); --k; }
  }
  if (i >= 0) n = push_cigar(cigar, n, 2, i + 1);
  if (k >= 0) n = push_cigar(cigar, n, 1, k + 1);
  for (int a = 0; a < n >> 1; ++a) { uint32_t tmp = cigar[a]; cigar[a] = cigar[n - 1 - a]; cigar[n - 1 - a] = tmp; }
  return n;
}

// Final alignment record (traceback + finalize, bwa/src/aux.cu:2874-3059).
struct Aln {
  int64_t pos, rb, re, cigar_off, core_off;
  int32_t rid, is_rev, flag, score, nw_score, nm, qb, qe, n_cigar, n_core;
};

// Cigar words an alignment may need: the global-alignment CIGAR (at most
// lq + rlen operations) followed by the clipped reference-format copy.
// An ungapped alignment (w = 0) needs at most three operations (clip, match,
// clip) and one core operation: eight words, so that neighbouring alignments
// share cache lines instead of each writing a few bytes into its own.
G3A_HD int64_t aln_cigar_cap(int lq, int64_t rlen, int w) { return w == 0 ? 8 : 2 * ((int64_t)lq + rlen) + 4; }
// Offset of the core CIGAR inside an alignment's words.
G3A_HD int64_t aln_core_at(int lq, int64_t rlen, int w) { return w == 0 ? 4 : (int64_t)lq + rlen + 4; }
// Direction bytes of the global alignment.
G3A_HD int64_t aln_z_size(int lq, int64_t rlen, int w) {
  if (w == 0) return 0;
  return (int64_t)z_stride(lq, w) * rlen;
}

// traceback (aux.cu:2874-2949) and finalize (aux.cu:2952-3059) of the ai-th
// kept region of a read of length l_seq.  `read` is the read, `cig` has
// aln_cigar_cap() words (the core CIGAR at aln_core_at()), z has
// aln_z_size() bytes, eh_h/eh_e lq+1 entries.
// pre_nw: when not null, the direction bytes and the score (*pre_nw) of the
// global alignment are already computed (k_tb_wave_dp) and only the walk runs.
template <class EH>
G3A_HD void trace_and_finalize(const DevIndex &x, const Opt &o, const uint8_t *read, int l_seq, const Region &g,
                               int ai, EH eh, uint8_t *z, uint32_t *cig, int64_t cig_base, Aln &a,
                               const int32_t *pre_nw = nullptr) {
  const int64_t l_pac = x.l_pac;
  const int lq = g.qe - g.qb;
  const int64_t rlen = g.re - g.rb;
  const uint8_t *query = read + g.qb;
  const int64_t rb = g.rb;
  const int w = global_bw(o, lq, rlen, g.score);
  const int64_t core_at = aln_core_at(lq, rlen, w);  // core CIGAR offset inside this record
  uint32_t *core = cig + core_at;
  int n_core = 0, nw = 0;
  TextReader tr(x.pac, l_pac);
  // NM as the reference's traceback counts it (aux.cu:2910-2941)
  int n_mm = 0, n_gap = 0;
  if (w == 0) {
    // one ungapped pass: score and mismatches together
    core[0] = (uint32_t)(lq << 4);
    n_core = 1;
    ByteReader qr(query);
    for (int i = 0; i < lq; ++i) {
      const int tb = tr(rb + i), qc = qr(i);
      nw += score_row(o, tb)(qc);
      n_mm += qc != tb;
    }
  } else {
    // the band revisits each query byte on 2w+1 rows: plain loads (cached)
    nw = pre_nw ? *pre_nw
                : global_dp(o, lq, [&](int j) { return (int)query[j]; }, (int)rlen, [&](int i) { return tr(rb + i); },
                            w, eh, z);
    n_core = global_walk(lq, (int)rlen, w, z, core);
    TextReader trn(x.pac, l_pac);
    ByteReader qn(query);
    int xq = 0;
    int64_t yr = 0;
    for (int k = 0; k < n_core; ++k) {
      int op = core[k] & 0xf, ln = (int)(core[k] >> 4);
      if (op == 0) {
        for (int i = 0; i < ln; ++i) n_mm += qn(xq + i) != trn(rb + yr + i);
        xq += ln;
        yr += ln;
      } else if (op == 2) {
        if (k > 0 && k < n_core - 1) n_gap += ln;
        yr += ln;
      } else if (op == 1) {
        xq += ln;
        n_gap += ln;
      }
    }
  }
  // finalize
  const int rev_strand = g.rb >= l_pac;
  int is_rev;
  int64_t pos = depos(x, !rev_strand ? g.rb : g.re - 1, is_rev);
  const uint32_t *cg = core;
  int nc = n_core;
  if (nc > 0) {
    if ((cg[0] & 0xf) == 2) { pos += cg[0] >> 4; --nc; cg = cg + 1; }
    else if ((cg[nc - 1] & 0xf) == 2) --nc;
  }
  int nf = 0;
  if (g.qb != 0 || g.qe != l_seq) {
    int clip5 = is_rev ? l_seq - g.qe : g.qb, clip3 = is_rev ? g.qb : l_seq - g.qe;
    if (clip5) cig[nf++] = (uint32_t)(clip5 << 4 | 3);
    for (int k = 0; k < nc; ++k) cig[nf++] = cg[k];
    if (clip3) cig[nf++] = (uint32_t)(clip3 << 4 | 3);
  } else {
    // the squeezed CIGAR is kept only when clipping is added (aux.cu:2985-3022)
    for (int k = 0; k < n_core; ++k) cig[nf++] = core[k];
  }
  a.rid = pos2rid(x, pos);
  a.pos = pos - x.ctg_off[a.rid];
  a.is_rev = is_rev;
  int flag = 0;
  if (g.secondary >= 0) flag |= 0x100;
  else if (ai > 0) flag |= 0x800;
  if (g.rid < 0) flag |= 0x4;
  if (is_rev) flag |= 0x10;
  a.flag = flag;
  a.score = g.score;
  a.nw_score = nw;
  a.nm = n_mm + n_gap;
  a.qb = g.qb;
  a.qe = g.qe;
  a.rb = g.rb;
  a.re = g.re;
  a.n_cigar = nf;
  a.n_core = n_core;
  a.cigar_off = cig_base;
  a.core_off = cig_base + core_at;
}

}  // namespace g3a
