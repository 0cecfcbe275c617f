// The learned-index form on the host, through the kernels' own code
// (g3a_dev.h, g3a_learned_fit.h compiled for the CPU):
//   - full SA from the sample walks against the LF-walk SA of every row;
//   - every row's key against the text read symbol by symbol; keys sorted;
//   - the model trained as the build kernel trains it (li_fit_leaf per
//     leaf), and its largest error over the distinct keys;
//   - li_lower_bound on every distinct key and on random keys against
//     std::lower_bound; li_range on text and random j-mers against the
//     FM-index's interval from forward extensions;
//   - the same lookups with a model whose weights are zeroed (windows miss:
//     every answer given must still be exact);
//   - seeding (rounds 1-3) of simulated reads with and without the learned
//     form: identical (k, s, m, n) lists.
// Prints "ok key=value ..." or the first failure.  Built and run by
// tests/test_learned.py.
#include "g3sa_amd/csrc/g3a_dev.h"
#include "g3sa_amd/csrc/g3a_index.h"
#include "g3sa_amd/csrc/g3a_layout.h"
#include "g3sa_amd/csrc/g3a_learned_fit.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

using namespace g3a;

static int fail(const char *what, long long a, long long b, long long c) {
  std::printf("FAIL %s %lld %lld %lld\n", what, a, b, c);
  return 1;
}

struct Model {
  std::vector<uint32_t> saf, root, leaf_a, shortt;
  std::vector<uint64_t> keys;
  std::vector<LiLeaf> leaf;
  std::vector<LiKnot> knot;
};

// FM interval of the j-symbol string W by forward extensions from its
// first symbol
static void fm_range(const DevIndex &x, uint64_t W, int j, uint32_t &k, uint32_t &s) {
  IntvN iv;
  intv_init(x, iv, (int)(W >> 62));
  for (int t = 1; t < j && iv.s > 0; ++t) forward_ext(x, iv, (int)(W >> (62 - 2 * t)) & 3);
  k = iv.k;
  s = iv.s;
}

struct SeedOut {
  std::vector<Intv> iv;
  int64_t steps = 0;
};

// E, P: the list entries and interval arithmetic of the narrow layout
// (LepNarrow, uint32_t) or of texts of 2^32 rows or more (LepWide, int64_t)
template <class E = LepNarrow, class P = uint32_t>
static void seed_read(const DevIndex &x, const Opt &o, const std::vector<uint8_t> &q, SeedOut &out,
                      int64_t *lc) {
  std::vector<Intv> slot(4096);
  std::vector<E> la(400), lf(400), lep(400), sd(400);
  IntvSink sink{slot.data(), 4096, 0, 0, o.max_occ};
  const int len = (int)q.size();
  GlobalBuf<E> buf{GlobalList<E>{la.data()}, GlobalList<E>{lf.data()}, 400, 400};
  if (!smem_round1(x, o, q.data(), len, sink, buf, &out.steps, lc)) {
    sink.n = 0;
    smem_round1_per_end(x, o, q.data(), len, sink, &out.steps);
  }
  const int nsm = sink.n;
  for (int k = 0; k < nsm; ++k)
    reseed_one(x, o, q.data(), len, slot[k], GlobalList<E>{lep.data()}, GlobalList<E>{sd.data()}, 400, 400, sink,
               &out.steps);
  reseed_round3<P>(x, o, q.data(), len, sink, &out.steps, lc);
  out.iv.assign(slot.begin(), slot.begin() + sink.n);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const int n_reads = argc > 2 ? std::atoi(argv[2]) : 2000;
  HostIndex h;
  std::string err;
  if (!load_index(argv[1], h, err)) {
    std::printf("load failed: %s\n", err.c_str());
    return 1;
  }
  unpack_bwt(h);
  h.pac.resize(h.pac.size() + 32, 0);
  DevIndex x;
  std::memset(&x, 0, sizeof x);
  std::vector<uint32_t> wide, narrow, sb;
  if (!build_occ_lines(h, wide, x.super) || !build_occ_narrow(h, narrow, sb)) {
    std::printf("layout failed\n");
    return 1;
  }
  x.occ = wide.data();
  x.occn = narrow.data();
  x.sbc = sb.data();
  x.primary = (int64_t)h.primary;
  x.n = (int64_t)h.seq_len;
  x.l_pac = h.l_pac;
  x.pac = h.pac.data();
  build_counts(h, x.count);
  if (h.sa_width == 4) x.sa32 = (const uint32_t *)h.sa.data();
  else x.sa64 = (const uint64_t *)h.sa.data();
  x.sa_shift = log2_exact(h.sa_intv);
  const int64_t n = x.n, N = n + 1;

  // full SA from the sample walks
  Model M;
  M.saf.assign((size_t)N, 0xffffffffu);
  const int64_t n_samples = (int64_t)(h.sa.size() / h.sa_width);
  for (int64_t j = 0; j < n_samples; ++j)
    if (!li_walk_sample(x, j, M.saf.data())) return fail("walk", j, 0, 0);
  for (int64_t r = 0; r < N; ++r)
    if ((int64_t)M.saf[r] != sa_value_n(x, (uint32_t)r)) return fail("saf", r, M.saf[r], sa_value_n(x, (uint32_t)r));

  // keys against the text symbol by symbol
  auto sym = [&](int64_t p) -> uint64_t {
    if (p >= n) return 0;
    const int64_t f = p < x.l_pac ? p : n - 1 - p;
    const uint64_t c = (h.pac[(size_t)(f >> 2)] >> ((~f & 3) << 1)) & 3;
    return p < x.l_pac ? c : 3 - c;
  };
  M.keys.resize((size_t)N + 16);  // padded as on the device
  for (int64_t r = 0; r < N; ++r) {
    const int64_t p = M.saf[r];
    uint64_t w = 0;
    for (int t = 0; t < 32; ++t) w |= sym(p + t) << (62 - 2 * t);
    M.keys[r] = text_word(x, p);
    if (M.keys[r] != w) return fail("key", r, p, 0);
    if (r && M.keys[r - 1] > M.keys[r]) return fail("order", r, 0, 0);
    if (n - p < kLiShort) {
      M.shortt.push_back((uint32_t)r);
      M.shortt.push_back((uint32_t)(n - p));
    }
  }
  if ((int64_t)M.shortt.size() != 2 * std::min<int64_t>(N, kLiShort)) return fail("short", M.shortt.size(), 0, 0);
  while (M.shortt.size() < 2 * kLiShort) M.shortt.push_back(0xffffffffu);

  // root and leaves, as the build kernels do it
  const int S = 1 << kLiRootBits;
  M.root.assign(S + 1, 0);
  for (int s = 0, r = 0; s <= S; ++s) {
    while (r < N && (int64_t)(M.keys[r] >> (64 - kLiRootBits)) < s) ++r;
    M.root[s] = (uint32_t)r;
  }
  x.saf = M.saf.data();
  x.keys = M.keys.data();
  x.li_root = M.root.data();
  x.li_short = M.shortt.data();
  x.li_leaf_w = (uint32_t)((N + kLiLeaves - 1) / kLiLeaves);
  M.leaf_a.assign(kLiLeaves + 1, (uint32_t)N);
  for (int64_t r = N - 1; r >= 0; --r) M.leaf_a[li_route(x, M.keys[r])] = (uint32_t)r;
  for (int j = kLiLeaves - 1; j >= 0; --j) M.leaf_a[j] = std::min(M.leaf_a[j], M.leaf_a[j + 1]);
  for (int64_t r = 1; r < N; ++r)
    if (li_route(x, M.keys[r]) < li_route(x, M.keys[r - 1])) return fail("route", r, 0, 0);
  M.leaf.resize(kLiLeaves);
  uint32_t n_knots = 0, max_knots = 0;
  for (int j = 0; j < kLiLeaves; ++j) {
    LiLeaf &L = M.leaf[j];
    L.a = M.leaf_a[j];
    L.err = 0;
    L.k0 = n_knots;
    L.nk = li_fit_pass(M.keys.data(), M.leaf_a[j], M.leaf_a[j + 1], kLiFitEps, nullptr);
    n_knots += L.nk;
    max_knots = std::max(max_knots, L.nk);
  }
  M.knot.resize(n_knots + 1);
  for (int j = 0; j < kLiLeaves; ++j)
    if (li_fit_pass(M.keys.data(), M.leaf_a[j], M.leaf_a[j + 1], kLiFitEps, M.knot.data() + M.leaf[j].k0) !=
        M.leaf[j].nk)
      return fail("knot count", j, 0, 0);
  x.li_leaf = M.leaf.data();
  x.li_knot = M.knot.data();
  int64_t distinct = 0;
  for (int64_t r = 0; r < N; ++r) distinct += !(r && M.keys[r - 1] == M.keys[r]);
  int64_t esum = 0, emax = 0, over = 0;
  for (int64_t r = 0; r < N; ++r) {
    if (r && M.keys[r - 1] == M.keys[r]) continue;
    const uint32_t j = li_route(x, M.keys[r]);
    const int64_t e = std::llabs(li_predict(x.li_knot, M.leaf[j], M.keys[r]) - r);
    M.leaf[j].err = std::max<uint32_t>(M.leaf[j].err, (uint32_t)std::min<int64_t>(e, 0x7fffffff));
    esum += e;
    emax = std::max(emax, e);
    over += e + 2 > kLiWindowCap;
  }

  std::mt19937_64 rng(12345);
  int64_t rmiss_long = 0;  // range misses of strings of kKmer symbols or more
  auto check_lookups = [&](bool trained, int64_t &misses, int64_t &ranges, int64_t &range_miss) -> int {
    misses = ranges = range_miss = rmiss_long = 0;
    for (int64_t r = 0; r < N; ++r) {
      if (r && M.keys[r - 1] == M.keys[r]) continue;
      const int64_t lb = li_lower_bound(x, M.keys[r]);
      const LiLeaf &L = M.leaf[li_route(x, M.keys[r])];
      if (lb < 0) {
        ++misses;
        if (trained && L.err + 2 <= (uint32_t)kLiWindowCap) return fail("trained miss", r, L.err, 0);
      } else if (lb != r) {
        return fail("lb", r, lb, 0);
      }
    }
    for (int i = 0; i < 200000; ++i) {
      uint64_t K = rng();
      if (i & 1) K = M.keys[rng() % N] + (rng() % 3) - 1;  // near a text key
      const int64_t lb = li_lower_bound(x, K);
      const int64_t want = std::lower_bound(M.keys.begin(), M.keys.begin() + N, K) - M.keys.begin();
      if (lb >= 0 && lb != want) return fail("lb random", i, lb, want);
    }
    for (int i = 0; i < 200000; ++i) {
      const int j = 1 + (int)(rng() % 32);
      uint64_t W;
      if (i % 4 == 3) {
        W = rng();
      } else {
        const int64_t p = (int64_t)(rng() % (uint64_t)N);
        if (p + j > n) continue;
        W = text_word(x, p);
      }
      W &= ~0ULL << (64 - 2 * j);
      uint32_t k1, s1, k2, s2;
      fm_range(x, W, j, k1, s1);
      ++ranges;
      if (!li_range(x, W, j, k2, s2)) {
        ++range_miss;
        rmiss_long += j >= kKmer;
        continue;
      }
      if (s1 != s2 || (s1 && k1 != k2)) return fail("range", i, j, (long long)s1 - (long long)s2);
    }
    return 0;
  };
  int64_t miss_t, ranges_t, rmiss_t, miss_z, ranges_z, rmiss_z, miss_b, ranges_b, rmiss_b;
  // the model's windows alone (no k-mer table)
  if (check_lookups(true, miss_t, ranges_t, rmiss_t)) return 1;
  // with the k-mer table, as the product runs: strings of kKmer symbols or
  // more are searched inside their k-mer's interval and never miss
  std::vector<uint32_t> kmer_table((size_t)3 << (2 * kKmer));
  {
    std::vector<std::thread> th;
    for (int t = 0; t < 8; ++t)
      th.emplace_back([&, t] {
        for (uint32_t c = t; c < (1u << (2 * kKmer)); c += 8) kmer_interval(x, c, &kmer_table[3 * (size_t)c]);
      });
    for (auto &t : th) t.join();
  }
  x.kmer = kmer_table.data();
  if (check_lookups(true, miss_b, ranges_b, rmiss_b)) return 1;
  if (rmiss_long) return fail("bucketed miss", rmiss_long, 0, 0);

  // seeding of simulated reads, plain and learned
  Opt o;
  std::memset(&o, 0, sizeof o);
  o.a = 1; o.b = 4; o.o_del = o.o_ins = 6; o.e_del = o.e_ins = 1; o.w = 100; o.T = 30;
  o.pen_clip5 = o.pen_clip3 = 5; o.min_seed_len = 19; o.split_width = 10; o.max_occ = 500;
  o.max_chain_gap = 10000; o.max_mem_intv = 20; o.split_factor = 1.5f; o.mask_level = 0.5f;
  o.drop_ratio = 0.5f; o.mask_level_redun = 0.95f;
  DevIndex xp = x;
  xp.saf = nullptr;
  xp.keys = nullptr;
  xp.li_root = nullptr;
  xp.li_leaf = nullptr;
  xp.li_short = nullptr;
  // the layout of texts of 2^32 rows or more (32-byte lines, 64-bit
  // intervals, no k-mer table), run on this text: the seeding code that
  // such a text takes on the device, checked here against the narrow path
  DevIndex xw = xp;
  xw.occn = nullptr;
  xw.sbc = nullptr;
  xw.kmer = nullptr;
  int64_t lc[6] = {0, 0, 0, 0, 0, 0}, n_iv = 0, steps_p = 0, steps_l = 0;
  for (int i = 0; i < n_reads; ++i) {
    const int len = 70 + (int)(rng() % 181);
    std::vector<uint8_t> q((size_t)len);
    if (i % 10 == 9 || len >= n) {
      for (auto &b : q) b = (uint8_t)(rng() % 4);
    } else {
      const int64_t p = (int64_t)(rng() % (uint64_t)(n - len));
      for (int t = 0; t < len; ++t) q[t] = (uint8_t)sym(p + t);
      const int err_pm = i % 3 == 0 ? 0 : (i % 3 == 1 ? 10 : 50);  // per mille
      for (int t = 0; t < len; ++t) {
        if ((int)(rng() % 1000) < err_pm) q[t] = (uint8_t)((q[t] + 1 + rng() % 3) & 3);
        if (rng() % 1000 == 0) q[t] = 4;
      }
    }
    SeedOut a, b, w;
    seed_read(xp, o, q, a, nullptr);
    seed_read(x, o, q, b, lc);
    seed_read<LepWide, int64_t>(xw, o, q, w, nullptr);
    if (a.iv.size() != w.iv.size()) return fail("wide seed count", i, a.iv.size(), w.iv.size());
    for (size_t k = 0; k < a.iv.size(); ++k) {
      const Intv &u = a.iv[k], &v = w.iv[k];
      if (u.k != v.k || u.s != v.s || u.m != v.m || u.n != v.n) return fail("wide seed", i, (long long)k, u.m);
    }
    steps_p += a.steps;
    steps_l += b.steps;
    if (a.iv.size() != b.iv.size()) return fail("seed count", i, a.iv.size(), b.iv.size());
    for (size_t k = 0; k < a.iv.size(); ++k) {
      const Intv &u = a.iv[k], &v = b.iv[k];
      if (u.k != v.k || u.s != v.s || u.m != v.m || u.n != v.n) return fail("seed", i, (long long)k, u.m);
    }
    n_iv += (int64_t)a.iv.size();
  }

  // a model with zeroed weights: windows of +-2 around the leaf's first
  // row mostly miss, and every answer given must still be exact
  for (auto &K : M.knot) K.y = K.s = 0.f;
  for (auto &L : M.leaf) L.err = 0;
  if (check_lookups(false, miss_z, ranges_z, rmiss_z)) return 1;
  if (rmiss_long) return fail("zeroed bucketed miss", rmiss_long, 0, 0);

  std::printf("ok rows=%lld distinct=%lld max_err=%lld mean_err=%.2f over_window=%lld lb_miss=%lld "
              "range_checked=%lld range_miss=%lld zeroed_lb_miss=%lld zeroed_range_miss=%lld reads=%d "
              "intervals=%lld lookups=%lld lookup_miss=%lld text_runs=%lld run_bases=%lld steps_plain=%lld "
              "steps_learned=%lld model_narrowed=%lld model_missed=%lld bucketed_range_miss=%lld knots=%u "
              "max_knots=%u\n",
              (long long)N, (long long)distinct, (long long)emax, distinct ? (double)esum / distinct : 0.0,
              (long long)over, (long long)miss_t, (long long)ranges_t, (long long)rmiss_t, (long long)miss_z,
              (long long)rmiss_z, n_reads, (long long)n_iv, (long long)lc[0], (long long)lc[1], (long long)lc[2],
              (long long)lc[3], (long long)steps_p, (long long)steps_l, (long long)lc[4], (long long)lc[5],
              (long long)rmiss_b, n_knots, max_knots);
  return 0;
}
