// The lane-per-task extension engine (g3a_ext.h, compiled for the host) on
// random tasks: the 32-bit engine at strip heights 16 and 32 against the
// oracle's banded extension (orc_ksw_extend), and the packed engine (two
// strip rows per register) against the 32-bit one, field by field.  Prints
// the counts; exits non-zero on any mismatch.  Built and run by
// tests/test_lane_engine.py.
#include "g3sa_amd/csrc/g3a_ext.h"
extern "C" {
#include "oracle/oracle.h"
}
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <utility>
#include <vector>
using namespace g3a;

int main(int argc, char **argv) {
  int n = argc > 1 ? atoi(argv[1]) : 100000;
  uint64_t seed = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
  std::mt19937_64 rng(seed);
  auto U = [&](int lo, int hi) { return (int)(lo + rng() % (uint64_t)(hi - lo + 1)); };
  Opt o;
  memset(&o, 0, sizeof o);
  o.a = 1; o.b = 4; o.o_del = o.o_ins = 6; o.e_del = o.e_ins = 1; o.w = 100;
  int k = 0;
  for (int i = 0; i < 4; ++i) { for (int j = 0; j < 4; ++j) o.mat[k++] = (int8_t)(i == j ? 1 : -4); o.mat[k++] = -1; }
  for (int j = 0; j < 5; ++j) o.mat[k++] = -1;
  std::vector<uint32_t> rb(kExtQCap), wt(6 * 4);
  long bad = 0, tested = 0, skipped = 0, pruned = 0, h0_none = 0, h0_cell = 0;
  long long cells_o = 0, cells_l = 0;
  for (int it = 0; it < n; ++it) {
    if (rng() % 3 == 0) { o.o_del = U(3, 8); o.e_del = U(1, 2); o.o_ins = U(3, 8); o.e_ins = U(1, 2); }
    else { o.o_del = o.o_ins = 6; o.e_del = o.e_ins = 1; }
    int qlen = rng() % 4 == 0 ? U(1, 320) : U(1, 150);
    int mode = U(0, 10);
    int ww = mode == 9 ? U(1, 40) : 100;
    o.w = ww;
    int tlen = U(0, qlen + 2 * ww + 20);
    int h0 = mode == 8 ? U(0, 3) : U(1, 320);
    int end_bonus = 5;
    std::vector<uint8_t> q(qlen), t(tlen > 0 ? tlen : 1);
    for (auto &c : q) c = U(0, 3);
    // target: mutated query with indels, then random
    int sub = U(0, 20), pos = 0;
    for (int i = 0; i < tlen; ++i) {
      if (pos < qlen && mode < 7) {
        int r = U(0, 999);
        if (r < sub * 5) t[i] = U(0, 3);
        else t[i] = q[pos];
        if (r >= 990) ++pos;          // deletion in target
        else if (r >= 985) { --i; ++pos; continue; }
        ++pos;
      } else t[i] = U(0, 3);
      if (U(0, 999) == 0) t[i] = 4;
    }
    if (U(0, 20) == 0) q[U(0, qlen - 1)] = 4;
    if (mode == 10) {
      // the target is the query with a mismatch every p-th base (p odd, so
      // coprime to 32): the diagonal comes back to exactly h0 every p rows
      // and never exceeds it, so over enough rows every residue of the rows
      // mod 32 holds a cell equal to h0 (the residue rule's "a cell" case)
      qlen = U(100, 320);
      q.resize(qlen);
      for (auto &c : q) c = U(0, 3);
      tlen = qlen;
      t.assign(q.begin(), q.end());
      const int per = 2 * U(1, 4) + 1, ph = U(0, per - 1);
      for (int i = ph; i < tlen; i += per) t[i] = (q[i] + 1 + U(0, 2)) & 3;
      o.o_del = o.o_ins = 6; o.e_del = o.e_ins = 1; o.w = 100;
      h0 = U(30, 200);
    }
    int w = ext_band_w(o, qlen, end_bonus);
    if (!ext_lane_ok(o, qlen, w, h0)) { ++skipped; continue; }
    int qle, tle;
    int64_t c_o = 0;
    int sc = orc_ksw_extend(qlen, q.data(), tlen, t.data(), o.mat, o.o_del, o.e_del, o.o_ins, o.e_ins, o.w, end_bonus,
                            h0, &qle, &tle, &c_o);
    int qmax = U(qlen, kExtQCap), wmin = U(0, w), wmax = w + U(0, 60);
    if (U(0, 1)) { qmax = qlen; wmin = w; wmax = w; }
    auto qb = [&](int j) { return (int)q[j]; };
    auto tb = [&](int i) { return (int)t[i]; };
    const bool sym = o.o_del + o.e_del == o.o_ins + o.e_ins;
    ExtLaneOut a = sym ? ext_lane<16, true>(o, qlen, qb, tlen, tb, w, h0, qmax, wmin, wmax, rb.data(), 1, wt.data(), 2)
                       : ext_lane<16, false>(o, qlen, qb, tlen, tb, w, h0, qmax, wmin, wmax, rb.data(), 1, wt.data(), 2);
    ExtLaneOut b = ext_lane<32, false>(o, qlen, qb, tlen, tb, w, h0, qmax, wmin, wmax, rb.data(), 1, wt.data(), 4);
    ExtLaneOut pa = sym ? ext_lane<16, true, true>(o, qlen, qb, tlen, tb, w, h0, qmax, wmin, wmax, rb.data(), 1, wt.data(), 2)
                        : ext_lane<16, false, true>(o, qlen, qb, tlen, tb, w, h0, qmax, wmin, wmax, rb.data(), 1, wt.data(), 2);
    ExtLaneOut pb = ext_lane<32, false, true>(o, qlen, qb, tlen, tb, w, h0, qmax, wmin, wmax, rb.data(), 1, wt.data(), 4);
    for (auto pr : {std::make_pair(&a, &pa), std::make_pair(&b, &pb)}) {
      const ExtLaneOut &x = *pr.first, &y = *pr.second;
      if (x.score != y.score || x.qle != y.qle || x.tle != y.tle || x.cells != y.cells) {
        if (bad < 10)
          fprintf(stderr, "PACKED MISMATCH it=%d qlen=%d tlen=%d w=%d h0=%d: (%d,%d,%d) vs packed (%d,%d,%d)\n", it,
                  qlen, tlen, w, h0, x.score, x.qle, x.tle, y.score, y.qle, y.tle);
        ++bad;
      }
    }
    ++tested;
    cells_o += c_o; cells_l += a.cells;
    if (a.cells < c_o) ++pruned;
    if (sc == h0) ++(qle ? h0_cell : h0_none);  // the residue rule: a cell, or none
    for (auto *x : {&a, &b}) {
      if (x->score != sc || x->qle != qle || x->tle != tle) {
        if (bad < 10)
          fprintf(stderr, "MISMATCH it=%d qlen=%d tlen=%d w=%d h0=%d: oracle (%d,%d,%d) lane%s (%d,%d,%d)\n", it, qlen,
                  tlen, w, h0, sc, qle, tle, x == &a ? "16" : "32", x->score, x->qle, x->tle);
        ++bad;
      }
    }
  }
  printf("tested %ld skipped %ld bad %ld pruned %ld h0_cell %ld h0_none %ld cells oracle %lld lane16 %lld (%.2f)\n",
         tested, skipped, bad, pruned, h0_cell, h0_none, cells_o, cells_l, (double)cells_l / cells_o);
  return bad != 0;
}
