// Banded local extension, lane-per-task engine (ksw_extend_warp2,
// bwa/src/ksw.cu:814-1034, as its cell set and result are restated in the
// oracle's orc_ksw_extend).
//
// Each lane of a wavefront runs one extension task.  The DP is swept in
// strips of R target rows: the strip's current column of H (and the F value
// of each row) lives in registers, so a cell costs a dozen VALU operations
// and no memory access; the row above the strip (H, E and the query code of
// each column, packed in one 32-bit word) is read and rewritten once per
// column from a per-lane row buffer laid out column-major across the
// wavefront, so a column is one coalesced 256-byte access.
//
// The engine applies every rule of the reference's cell set: the band's left
// edge, the column limit 32*(i/32+1)+w+1, the corner rule and the quirk rows
// (only in the few columns of a strip where one of them can apply to some
// lane; the other columns run without any test).  It admits every task
// whose scores fit its 12-bit cells.  When no cell exceeds h0, the
// reference's position is decided by its 32-lane residue rule: no cell
// unless every residue of the rows mod 32 holds a cell equal to h0, and then
// the first such row and its last such column.  While the best is h0 the
// engine records, in the rare columns whose maximum equals h0, which rows
// hold h0 (cells outside the target or query, or outside the band, never
// do: they are below every real cell or 0).
//
// Exact pruning: after each strip the lane stops once no later cell can
// reach the current best.  A later cell's value comes from a positive cell of
// the strip's last row (through the diagonal or E, both bounded by that
// cell's H) gaining at most max(mat) per remaining query column, from the
// column -1 boundary max(h0 - o_del - e_del*i, 0) of a later row, or from a
// later quirk row's injected diagonal; when all of these bounds are below the
// best, no later cell can equal or exceed it, so neither the score nor its
// position can change.
#pragma once
#include <type_traits>

#include "g3a_dev.h"

namespace g3a {

constexpr int kExtQCap = kMaxReadLen;  // row-buffer columns per lane

// max of three signed ints as one VALU instruction on the device (the
// compiler otherwise splits a signed/unsigned pair of maxima)
G3A_HD int max3i(int a, int b, int c) {
#if defined(__HIP_DEVICE_COMPILE__)
  int d;
  asm("v_max3_i32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
#else
  const int t = a > b ? a : b;
  return t > c ? t : c;
#endif
}
constexpr int kExtPad = 5;             // query/target code of padding cells

// Two 16-bit lanes of one 32-bit register: VOP3P packed integer
// instructions on the device (v_pk_add_u16, v_pk_max_i16, v_pk_sub_u16 with
// clamp, ...), a pair of int16 on the host.  Signed unless named otherwise.
#if defined(__HIP_DEVICE_COMPILE__)
typedef short g3a_s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short g3a_u16x2 __attribute__((ext_vector_type(2)));
struct P2 {
  g3a_s16x2 v;
};
G3A_HD P2 p2(int lo, int hi) {
  P2 r;
  r.v = (g3a_s16x2){(short)lo, (short)hi};
  return r;
}
G3A_HD int p2_lo(P2 a) { return a.v.x; }
G3A_HD int p2_hi(P2 a) { return a.v.y; }
G3A_HD P2 p2_add(P2 a, P2 b) { P2 r; r.v = a.v + b.v; return r; }
G3A_HD P2 p2_sub(P2 a, P2 b) { P2 r; r.v = a.v - b.v; return r; }
G3A_HD P2 p2_max(P2 a, P2 b) { P2 r; r.v = __builtin_elementwise_max(a.v, b.v); return r; }
G3A_HD P2 p2_min(P2 a, P2 b) { P2 r; r.v = __builtin_elementwise_min(a.v, b.v); return r; }
// max(a - b, 0) for a, b >= 0 (unsigned saturating subtract)
G3A_HD P2 p2_subsat(P2 a, P2 b) {
  const g3a_u16x2 ua = __builtin_bit_cast(g3a_u16x2, a.v), ub = __builtin_bit_cast(g3a_u16x2, b.v);
  P2 r;
  r.v = __builtin_bit_cast(g3a_s16x2, __builtin_elementwise_sub_sat(ua, ub));
  return r;
}
G3A_HD P2 p2_shl3(P2 a) {
  P2 r;
  r.v = __builtin_bit_cast(g3a_s16x2, __builtin_bit_cast(g3a_u16x2, a.v) << (g3a_u16x2){3, 3});
  return r;
}
// the signed 4-bit field q (0..3) of each 16-bit half of x
G3A_HD P2 p2_nibble(uint32_t x, int q) {
  const unsigned short sh = (unsigned short)(12 - 4 * q);
  const g3a_u16x2 u = __builtin_bit_cast(g3a_u16x2, x) << (g3a_u16x2){sh, sh};
  P2 r;
  r.v = __builtin_bit_cast(g3a_s16x2, u) >> (g3a_s16x2){12, 12};
  return r;
}
#else
struct P2 {
  int16_t lo, hi;
};
G3A_HD P2 p2(int lo, int hi) { return P2{(int16_t)lo, (int16_t)hi}; }
G3A_HD int p2_lo(P2 a) { return a.lo; }
G3A_HD int p2_hi(P2 a) { return a.hi; }
G3A_HD P2 p2_add(P2 a, P2 b) { return p2(a.lo + b.lo, a.hi + b.hi); }
G3A_HD P2 p2_sub(P2 a, P2 b) { return p2(a.lo - b.lo, a.hi - b.hi); }
G3A_HD P2 p2_max(P2 a, P2 b) { return p2(a.lo > b.lo ? a.lo : b.lo, a.hi > b.hi ? a.hi : b.hi); }
G3A_HD P2 p2_min(P2 a, P2 b) { return p2(a.lo < b.lo ? a.lo : b.lo, a.hi < b.hi ? a.hi : b.hi); }
G3A_HD P2 p2_subsat(P2 a, P2 b) {
  const int l = (uint16_t)a.lo - (uint16_t)b.lo, h = (uint16_t)a.hi - (uint16_t)b.hi;
  return p2(l > 0 ? l : 0, h > 0 ? h : 0);
}
G3A_HD P2 p2_shl3(P2 a) { return p2((int16_t)(uint16_t)((uint16_t)a.lo << 3), (int16_t)(uint16_t)((uint16_t)a.hi << 3)); }
G3A_HD P2 p2_nibble(uint32_t x, int q) {
  const int lo = (int)((x >> (4 * q)) & 15), hi = (int)((x >> (16 + 4 * q)) & 15);
  return p2(lo >= 8 ? lo - 16 : lo, hi >= 8 ? hi - 16 : hi);
}
#endif

// Band width after clipping by max_ins/max_del (ksw.cu:840-851).
G3A_HD int ext_band_w(const Opt &o, int qlen, int end_bonus) {
  int max_ = 0;
  for (int i = 0; i < 25; ++i) max_ = max_ > o.mat[i] ? max_ : o.mat[i];
  int w = o.w;
  int max_ins = (int)((double)(qlen * max_ + end_bonus - o.o_ins) / o.e_ins + 1.);
  max_ins = max_ins > 1 ? max_ins : 1;
  w = w < max_ins ? w : max_ins;
  int max_del = (int)((double)(qlen * max_ + end_bonus - o.o_del) / o.e_del + 1.);
  max_del = max_del > 1 ? max_del : 1;
  w = w < max_del ? w : max_del;
  return w;
}

G3A_HD int mat_max(const Opt &o) {
  int m = 0;
  for (int i = 0; i < 25; ++i) m = m > o.mat[i] ? m : o.mat[i];
  return m;
}

// Scoring admitted by the engine: every score in [-8, 7] (4-bit fields).
G3A_HD bool ext_lane_opt_ok(const Opt &o) {
  for (int i = 0; i < 25; ++i)
    if (o.mat[i] < -8 || o.mat[i] > 7) return false;
  return o.e_del >= 0 && o.e_ins >= 0 && o.o_del >= 0 && o.o_ins >= 0;
}

// Can the engine compute this task exactly?  (Scores must fit its 12-bit
// cells; every rule of the cell set is applied in the general columns; h0 is
// at least 1, so that no cell outside the cell set, all 0, can equal it.)
G3A_HD bool ext_lane_ok(const Opt &o, int qlen, int w, int h0) {
  (void)w;
  if (qlen <= 0 || qlen > kExtQCap || h0 < 1) return false;
  return h0 + mat_max(o) * qlen <= 4095;
}

struct ExtLaneOut {
  int score, qle, tle;
  int64_t cells;
};

// One task on one lane.
//   qb(j): query code of column j < qlen; tb(i): target code of row i < tlen
//   qmax >= qlen, w_min <= w <= w_max: bounds shared by the lanes of a
//     wavefront (they keep the column loop and the memory accesses uniform;
//     any looser values give the same result)
//   rowbuf: column j of this lane at rowbuf[j * rs]
//   wt: score words of query code c at wt[c * cs + k], k < R/8
//   SYM: o_del + e_del == o_ins + e_ins (one M - o - e per cell)
//   PK: the plain columns run with rows r and r + R/2 of a strip packed in
//     one register (16-bit lanes), the lower half one column behind the
//     upper, so that each packed operation computes two cells
template <int R, bool SYM, bool PK = false, class QB, class TB>
G3A_HD ExtLaneOut ext_lane(const Opt &o, int qlen, QB qb, int tlen, TB tb, int w, int h0, int qmax, int w_min,
                           int w_max, uint32_t *rowbuf, int rs, uint32_t *wt, int cs) {
  static_assert(R == 16 || R == 32, "strip height");
  // SYM: the caller guarantees o_del + e_del == o_ins + e_ins
  constexpr int NW = R / 8;
  const int oe_del = o.o_del + o.e_del, oe_ins = o.o_ins + o.e_ins;
  const int ms = mat_max(o);
  // signed 4-bit scores of target code t against query codes 0..4
  uint32_t rwt[5];
  for (int t = 0; t < 5; ++t) {
    uint32_t v = 0;
    for (int c = 0; c < 5; ++c) v |= (uint32_t)(o.mat[t * 5 + c] & 15) << (4 * c);
    rwt[t] = v;
  }
  const uint32_t rw0 = rwt[0], rw1 = rwt[1], rw2 = rwt[2], rw3 = rwt[3], rw4 = rwt[4];
  // row -1: H(-1, j) is the diagonal input of cell (0, j+1), E(-1, j) = 0
  for (int j = 0; j < qmax; ++j) {
    int hv = h0 - oe_ins - j * o.e_ins;
    hv = hv > 0 ? hv : 0;
    const uint32_t qc = j < qlen ? (uint32_t)qb(j) : (uint32_t)kExtPad;
    rowbuf[(int64_t)j * rs] = (uint32_t)hv | (qc << 24);
  }
  int best = h0, bi = -1, bj = -1;
  // while best == h0: the row residues (mod 32) holding a cell equal to h0,
  // the first such row and its last such column
  uint32_t h0_rows = 0;
  int h0_i = -1, h0_j = -1;
  auto h0_note = [&](int row, int col) {
    if (row >= tlen || col >= qlen) return;
    h0_rows |= 1u << (row & 31);
    if (h0_i < 0 || row < h0_i) {
      h0_i = row;
      h0_j = col;
    } else if (row == h0_i && col > h0_j) {
      h0_j = col;
    }
  };
  int64_t cells = 0;
  int i0 = 0;
  for (;;) {
    // signed 4-bit scores of the strip's rows against each query code;
    // padding rows and columns score -8, below any real score
    {
      uint32_t wl[6][NW];
#pragma unroll
      for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int k = 0; k < NW; ++k) wl[c][k] = c == 5 ? 0x88888888u : 0u;  // column padding: -8
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = i0 + r;
        const int tc = i < tlen ? tb(i) : kExtPad;
        uint32_t rw = tc == 0 ? rw0 : tc == 1 ? rw1 : tc == 2 ? rw2 : tc == 3 ? rw3 : tc == 4 ? rw4 : 0x88888u;
#pragma unroll
        for (int c = 0; c < 5; ++c) wl[c][r >> 3] |= ((rw >> (4 * c)) & 15u) << ((r & 7) * 4);
      }
#pragma unroll
      for (int c = 0; c < 6; ++c)
#pragma unroll
        for (int k = 0; k < NW; ++k) wt[c * cs + k] = wl[c][k];
    }
    const int jstart = i0 - w_max > 0 ? i0 - w_max : 0;
    const int jfull = i0 + R - 1 - w_min;  // columns below this need the band mask
    int hc[R], F[R];
    int htop;  // diagonal input of row i0 at the next column
    if (jstart == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int b = h0 - o.o_del - o.e_del * (i0 + r + 1);
        hc[r] = b > 0 ? b : 0;
      }
      const int b0 = h0 - o.o_del - o.e_del * i0;
      htop = i0 == 0 ? h0 : (b0 > 0 ? b0 : 0);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) hc[r] = 0;
      htop = (int)(rowbuf[(int64_t)(jstart - 1) * rs] & 0xfff);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) F[r] = 0;
    int u = 0;
    const int lim0 = w - i0;  // cell (i0+r, j) is in the band iff r <= j + lim0
    // the reference's other rules for this strip: the column limit of its
    // 32-row block, the quirk row at the block start (i0 = 32k, k >= 1) whose
    // cells right of i0+w+1 take an injected diagonal and no E, and the
    // corner rule (i+j <= 30 needs j <= i+w+1)
    const int jlim = 32 * (i0 / 32 + 1) + w + 1;
    const bool quirk = i0 >= 32 && (i0 & 31) == 0;
    // one column of the strip; kGeneral: some rule above or the band's left
    // edge may apply to a row of this column
    auto column = [&](int j, auto general_tag) {
      constexpr bool kGeneral = decltype(general_tag)::value;
      const uint32_t word = rowbuf[(int64_t)j * rs];
      const int Ht = (int)(word & 0xfff);
      int Et = (int)((word >> 12) & 0xfff);
      const uint32_t qc = word >> 24;
      uint32_t W[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) W[k] = wt[qc * cs + k];
      int hd0 = htop;
      if (kGeneral && quirk && j > i0 + w + 1) {
        const int v = h0 - oe_ins - (j - 1) * o.e_ins;
        hd0 = v > 0 ? v : 0;
        Et = 0;
      }
      // M of every row from the previous column's H (bottom-up, in place).
      // M = 0 when the diagonal H is 0; otherwise H + score.  Computed as
      // min(hd + score, 8 hd): for hd >= 1 the bound 8 hd >= hd + 7 never
      // binds, for hd = 0 it gives a value <= 0.  A negative M acts as 0:
      // H = max(M, E, F) and E, F >= 0, and M - gap < 0 in either case.
#pragma unroll
      for (int r = R - 1; r >= 1; --r) {
        const int hd = hc[r - 1];
        const int m = hd + ((int)(W[r >> 3] << (28 - (r & 7) * 4)) >> 28);
        hc[r] = m < (hd << 3) ? m : (hd << 3);
      }
      {
        const int m = hd0 + ((int)(W[0] << 28) >> 28);
        hc[0] = m < (hd0 << 3) ? m : (hd0 << 3);
        htop = Ht;
      }
      int e = Et, cmax = 0;
      const int lim = j + lim0;                  // band: r <= lim
      const bool in_lim = j <= jlim;             // column limit
      const int cb = 30 - i0 - j, cc = j - i0 - w - 1;  // corner: r > cb or r >= cc
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int m = hc[r], f = F[r];
        int h = max3i(m, e, f);
        const int td = m - oe_del, ti = SYM ? td : m - oe_ins;
        int e2 = e - o.e_del;
        e2 = e2 > td ? e2 : td;
        e2 = e2 > 0 ? e2 : 0;
        int f2 = f - o.e_ins;
        f2 = f2 > ti ? f2 : ti;
        f2 = f2 > 0 ? f2 : 0;
        if (kGeneral && !(r <= lim && in_lim && (r > cb || r >= cc))) {
          h = 0;
          f2 = 0;
          e2 = 0;
        }
        hc[r] = h;
        F[r] = f2;
        e = e2;
        if (r & 1) cmax = max3i(cmax, hc[r - 1], h);
      }
      rowbuf[(int64_t)j * rs] = (uint32_t)hc[R - 1] | ((uint32_t)e << 12) | (qc << 24);
      if (best == h0 && cmax == h0) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (hc[r] == h0) h0_note(i0 + r, j);
      }
      // the column's best cell: only when it can change the result, find the
      // smallest row holding the column maximum
      if (cmax > best || (cmax == best && bi >= i0)) {
        int rmin = R - 1;
#pragma unroll
        for (int r = R - 2; r >= 0; --r)
          if (hc[r] == cmax) rmin = r;
        const int ci = i0 + rmin;
        if (cmax > best || ci <= bi) {
          best = cmax;
          bi = ci;
          bj = j;
        }
      }
      const int hl = hc[R - 1], ub = hl + ms * (qlen - 1 - j);
      if (hl > 0 && ub > u) u = ub;
    };
    // The best cell of column j from its two halves' maxima: cmax_u over rows
    // [0, R/2) with the smallest such row rmin_u (computed when it could be
    // needed), cmax_l over the rest, lower(r) the value of row R/2 + r; and
    // the bound from its last row hl.  (The packed columns' bookkeeping.)
    auto finish = [&](int jj, int cmax_u, int rmin_u, int cmax_l, auto lower, int hl) {
      constexpr int H = R / 2;
      const int cmax = cmax_u > cmax_l ? cmax_u : cmax_l;
      if (cmax > best || (cmax == best && bi >= i0)) {
        int rmin = rmin_u;
        if (cmax_u != cmax) {
          rmin = R - 1;
#pragma unroll
          for (int r = H - 1; r >= 0; --r)
            if (lower(r) == cmax) rmin = H + r;
        }
        const int ci = i0 + rmin;
        if (cmax > best || ci <= bi) {
          best = cmax;
          bi = ci;
          bj = jj;
        }
      }
      const int ub = hl + ms * (qlen - 1 - jj);
      if (hl > 0 && ub > u) u = ub;
    };
    // one plain cell of row r (32-bit), as in column()
    auto plain_cell = [&](int r, int &e, int &cm) {
      const int m = hc[r], f = F[r];
      const int h = max3i(m, e, f);
      const int td = m - oe_del, ti = SYM ? td : m - oe_ins;
      int e2 = e - o.e_del;
      e2 = e2 > td ? e2 : td;
      e2 = e2 > 0 ? e2 : 0;
      int f2 = f - o.e_ins;
      f2 = f2 > ti ? f2 : ti;
      f2 = f2 > 0 ? f2 : 0;
      hc[r] = h;
      F[r] = f2;
      e = e2;
      cm = cm > h ? cm : h;
    };
    // Plain columns [c0, c1), c1 - c0 >= 2, with rows r and r + H (H = R/2)
    // in one register: the upper half at column c, the lower half at c - 1.
    // The lower half's diagonal and E inputs at row H come from the upper
    // half's last row one and zero columns earlier; both chains (E down the
    // rows, F along the columns) keep their order in each half.
    auto packed_run = [&](int c0, int c1) {
      constexpr int H = R / 2, NG = H / 4;
      const P2 oeP = p2(oe_del, oe_del), oiP = p2(oe_ins, oe_ins), edP = p2(o.e_del, o.e_del),
               eiP = p2(o.e_ins, o.e_ins);
      // prologue: the upper rows of column c0
      int hul = hc[H - 1];  // upper row H-1, one column before the lower half's column
      uint32_t wp[NW];
      uint32_t qcp;
      int e_mid, cmax_u = 0, rmin_u = R - 1;
      {
        const uint32_t word = rowbuf[(int64_t)c0 * rs];
        const int Ht = (int)(word & 0xfff), Et = (int)((word >> 12) & 0xfff);
        qcp = word >> 24;
#pragma unroll
        for (int k = 0; k < NW; ++k) wp[k] = wt[qcp * cs + k];
#pragma unroll
        for (int r = H - 1; r >= 1; --r) {
          const int hd = hc[r - 1];
          const int m = hd + ((int)(wp[r >> 3] << (28 - (r & 7) * 4)) >> 28);
          hc[r] = m < (hd << 3) ? m : (hd << 3);
        }
        {
          const int m = htop + ((int)(wp[0] << 28) >> 28);
          hc[0] = m < (htop << 3) ? m : (htop << 3);
          htop = Ht;
        }
        int e = Et;
#pragma unroll
        for (int r = 0; r < H; ++r) plain_cell(r, e, cmax_u);
        e_mid = e;
        if (best == h0 && cmax_u == h0) {
#pragma unroll
          for (int r = 0; r < H; ++r)
            if (hc[r] == h0) h0_note(i0 + r, c0);
        }
        if (cmax_u > best || (cmax_u == best && bi >= i0)) {
#pragma unroll
          for (int r = H - 1; r >= 0; --r)
            if (hc[r] == cmax_u) rmin_u = r;
        }
      }
      P2 hP[H], fP[H];
#pragma unroll
      for (int r = 0; r < H; ++r) {
        hP[r] = p2(hc[r], hc[r + H]);
        fP[r] = p2(F[r], F[r + H]);
      }
      for (int c = c0 + 1; c < c1; ++c) {
        const uint32_t word = rowbuf[(int64_t)c * rs];
        const int Ht = (int)(word & 0xfff), Et = (int)((word >> 12) & 0xfff);
        const uint32_t qc = word >> 24;
        uint32_t wc[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) wc[k] = wt[qc * cs + k];
        // 4-row groups: upper rows' fields (column c) in the low half, lower
        // rows' (column c-1) in the high half, as signed 4-bit values
        uint32_t X[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
          const int sh = 16 * (g & 1);
          const uint32_t lo = (wc[g >> 1] >> sh) & 0xffffu, hi = (wp[H / 8 + (g >> 1)] >> sh) & 0xffffu;
          X[g] = lo | hi << 16;
        }
        const int nxt = p2_lo(hP[H - 1]);
#pragma unroll
        for (int r = H - 1; r >= 0; --r) {
          const P2 hd = r ? hP[r - 1] : p2(htop, hul);
          const P2 m = p2_add(hd, p2_nibble(X[r >> 2], r & 3));
          hP[r] = p2_min(m, p2_shl3(hd));
        }
        htop = Ht;
        hul = nxt;
        P2 eP = p2(Et, e_mid), cm = p2(0, 0);
#pragma unroll
        for (int r = 0; r < H; ++r) {
          const P2 m = hP[r], f = fP[r];
          const P2 h = p2_max(p2_max(m, eP), f);
          const P2 td = p2_sub(m, oeP), ti = SYM ? td : p2_sub(m, oiP);
          eP = p2_max(p2_subsat(eP, edP), td);
          fP[r] = p2_max(p2_subsat(f, eiP), ti);
          hP[r] = h;
          cm = p2_max(cm, h);
        }
        e_mid = p2_lo(eP);
        if (best == h0 && (p2_lo(cm) == h0 || p2_hi(cm) == h0)) {
#pragma unroll
          for (int r = 0; r < H; ++r) {
            if (p2_lo(hP[r]) == h0) h0_note(i0 + r, c);
            if (p2_hi(hP[r]) == h0) h0_note(i0 + H + r, c - 1);
          }
        }
        // column c-1 is complete
        const int hl = p2_hi(hP[H - 1]);
        rowbuf[(int64_t)(c - 1) * rs] = (uint32_t)hl | ((uint32_t)p2_hi(eP) << 12) | (qcp << 24);
        finish(c - 1, cmax_u, rmin_u, p2_hi(cm), [&](int r) { return p2_hi(hP[r]); }, hl);
        // the upper rows of column c
        cmax_u = p2_lo(cm);
        rmin_u = R - 1;
        if (cmax_u > best || (cmax_u == best && bi >= i0)) {
#pragma unroll
          for (int r = H - 1; r >= 0; --r)
            if (p2_lo(hP[r]) == cmax_u) rmin_u = r;
        }
#pragma unroll
        for (int k = 0; k < NW; ++k) wp[k] = wc[k];
        qcp = qc;
      }
      // epilogue: the lower rows of column c1-1
#pragma unroll
      for (int r = 0; r < H; ++r) {
        hc[r] = p2_lo(hP[r]);
        hc[r + H] = p2_hi(hP[r]);
        F[r] = p2_lo(fP[r]);
        F[r + H] = p2_hi(fP[r]);
      }
#pragma unroll
      for (int r = R - 1; r >= H; --r) {
        const int hd = r > H ? hc[r - 1] : hul;
        const int m = hd + ((int)(wp[r >> 3] << (28 - (r & 7) * 4)) >> 28);
        hc[r] = m < (hd << 3) ? m : (hd << 3);
      }
      int e = e_mid, cmax_l = 0;
#pragma unroll
      for (int r = H; r < R; ++r) plain_cell(r, e, cmax_l);
      if (best == h0 && cmax_l == h0) {
#pragma unroll
        for (int r = H; r < R; ++r)
          if (hc[r] == h0) h0_note(i0 + r, c1 - 1);
      }
      rowbuf[(int64_t)(c1 - 1) * rs] = (uint32_t)hc[R - 1] | ((uint32_t)e << 12) | (qcp << 24);
      finish(c1 - 1, cmax_u, rmin_u, cmax_l, [&](int r) { return hc[H + r]; }, hc[R - 1]);
    };
    // columns [ju0, ju1) need no rule for any lane of the wavefront
    int ju0 = jfull > jstart ? jfull : jstart;
    if (i0 + w_min + 1 < 30 - i0 && ju0 < 31 - i0) ju0 = 31 - i0;
    int ju1 = 32 * (i0 / 32 + 1) + w_min + 2;
    if (quirk && i0 + w_min + 2 < ju1) ju1 = i0 + w_min + 2;
    if (ju1 > qmax) ju1 = qmax;
    if (ju0 > qmax) ju0 = qmax;
    if (ju1 < ju0) ju1 = ju0;
    int j = jstart;
    for (; j < ju0; ++j) column(j, std::true_type());
    if (PK && ju1 - j >= 2) {
      packed_run(j, ju1);
      j = ju1;
    }
    for (; j < ju1; ++j) column(j, std::false_type());
    for (; j < qmax; ++j) column(j, std::true_type());
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = i0 + r;
      if (i >= tlen) continue;
      const int lo = i - w > 0 ? i - w : 0;
      int hi = qlen - 1 < jlim ? qlen - 1 : jlim;
      if (hi < lo) continue;
      int n = hi - lo + 1;
      const int c0 = lo > i + w + 2 ? lo : i + w + 2, c1 = hi < 30 - i ? hi : 30 - i;  // corner cells
      if (c1 >= c0) n -= c1 - c0 + 1;
      cells += n;
    }
    i0 += R;
    if (i0 >= tlen) break;
    // bound on every later cell: paths through the strip's last row, the
    // column -1 boundary of later rows, and the next quirk row's injections
    const int b = h0 - o.o_del - o.e_del * i0;
    int ub = b > 0 ? b + ms * qlen : 0;
    const int iq = (i0 + 31) & ~31;
    if (iq + w + 2 <= qlen - 1) {
      const int dq = h0 - oe_ins - (iq + w + 1) * o.e_ins;
      const int bq = dq > 0 ? dq + ms * (qlen - iq - w - 2) : 0;
      ub = ub > bq ? ub : bq;
    }
    if ((u > ub ? u : ub) < best) break;
  }
  if (best == h0) {  // the residue rule (bi and bj are still -1)
    if (h0_rows == 0xffffffffu) {
      bi = h0_i;
      bj = h0_j;
    }
  }
  ExtLaneOut out;
  out.score = best;
  out.qle = bj + 1;
  out.tle = bi + 1;
  out.cells = cells;
  return out;
}

}  // namespace g3a
