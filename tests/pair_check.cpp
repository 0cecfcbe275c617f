// Host check of g3a::repeat_add (g3sa_amd/csrc/g3a_pair.h) against the loop
// of single double adds it replaces: random sums, increments and counts,
// ties at several binades, increments below half an ulp, increments larger
// than the sum, and whole variance sums over size histograms.  Prints
// "ok <cases>" or the first mismatch.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <random>

#include "g3sa_amd/csrc/g3a_pair.h"

static double loop_add(double s, double t, uint32_t c) {
  volatile double v = s;  // one rounded add per step
  for (uint32_t i = 0; i < c; ++i) v = v + t;
  return v;
}

static bool same(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

int main() {
  std::mt19937_64 rng(7);
  int64_t cases = 0;
  auto check = [&](double s, double t, uint32_t c) {
    const double a = g3a::repeat_add(s, t, c), b = loop_add(s, t, c);
    ++cases;
    if (!same(a, b)) {
      std::printf("mismatch s=%a t=%a c=%u: %a != %a\n", s, t, c, a, b);
      return false;
    }
    return true;
  };
  // ties and near-ties at binades where the ulp is 1, 2, 4 and 2^-10
  const double tops[] = {9007199254740992.0 /* 2^53 */, 18014398509481984.0, 36028797018963968.0, 8796093022208.0};
  for (double top : tops) {
    const double u = top / 9007199254740992.0 * 2.0;  // ulp in [top, 2 top)
    for (int ds = 0; ds < 8; ++ds)
      for (double tf : {0.5, 1.5, 2.5, 0.49999999999, 0.5000000001, 3.0, 0.25, 1.0})
        for (uint32_t c : {1u, 2u, 3u, 7u, 1000u, 100000u})
          if (!check(top + ds * u, tf * u, c)) return 1;
  }
  // random sums and increments
  std::uniform_real_distribution<double> U(0.0, 1.0);
  for (int i = 0; i < 30000; ++i) {
    const double s = (i % 5 == 0) ? 0.0 : std::ldexp(U(rng), (int)(rng() % 80) - 20);
    const double t = std::ldexp(U(rng), (int)(rng() % 80) - 20);
    const uint32_t c = (uint32_t)(1 + rng() % (i % 10 == 0 ? 100000 : 300));
    if (!check(s, t, c)) return 1;
  }
  // variance sums: ascending sizes with counts, squared deviations from a
  // non-integer mean, as the insert-size statistics add them
  for (int rep = 0; rep < 60; ++rep) {
    const int lo = 1 + (int)(rng() % 400), hi = lo + 1 + (int)(rng() % 3000);
    const double avg = lo + (hi - lo) * U(rng);
    double a = 0, b = 0;
    for (int v = lo; v <= hi; ++v) {
      const uint32_t c = (uint32_t)(rng() % 4 == 0 ? 0 : rng() % (rep % 6 ? 50 : 20000));
      const double d = (double)v - avg, sq = d * d;
      a = g3a::repeat_add(a, sq, c);
      b = loop_add(b, sq, c);
      ++cases;
      if (!same(a, b)) {
        std::printf("variance mismatch rep %d v %d: %a != %a\n", rep, v, a, b);
        return 1;
      }
    }
  }
  std::printf("ok %" PRId64 "\n", cases);
  return 0;
}
