// Host/device helpers of the paired-end statistics.
#pragma once

#include "g3a_dev.h"

#include <cmath>
#include <cstdint>

namespace g3a {

// The value of s after adding t (t >= 0) to it c times in double precision,
// every sum rounded to nearest-even as in a loop of c adds, without running
// the loop.  While the running sum stays in one binade [top/2, top) with ulp
// u, an add from a sum that is a multiple of u rounds t to the same multiple
// of u every time, except that a tie (t exactly half-way between two
// multiples) rounds to the even one: after one add in the binade the sum is
// even in units of u and every later add there takes the same even
// increment.  So after two single adds (the second one measures the
// increment exactly: s1 <= s2 <= 2 s1) the adds that cannot leave the binade
// are applied as one exact multiple, and the adds at the binade's edge run
// one by one.  An increment of 0 (t below half an ulp) leaves the sum fixed.
G3A_HD double repeat_add(double s, double t, uint32_t c) {
  while (c) {
    if (t == 0) return s;
    if (!(s >= t)) {  // the sum at least doubles
      s += t;
      --c;
      continue;
    }
    int e;
    frexp(s, &e);  // s in [2^(e-1), 2^e)
    const double top = ldexp(1.0, e), u = ldexp(1.0, e - 53);
    const double s1 = s + t;
    --c;
    if (s1 >= top || c == 0) {
      s = s1;
      continue;
    }
    const double s2 = s1 + t;
    --c;
    if (s2 >= top || c == 0) {
      s = s2;
      continue;
    }
    const double inc = s2 - s1;
    if (inc == 0) return s2;
    // an add from sum x stays exact-below top when x + inc + u/2 < top,
    // i.e. (top - x)/u >= inc/u + 1 in units of u
    const int64_t R = (int64_t)((top - s2) / u), I = (int64_t)(inc / u);
    int64_t k = R - I - 1 >= 0 ? (R - I - 1) / I + 1 : 0;
    if (k > (int64_t)c) k = (int64_t)c;
    s = s2 + (double)(k * I) * u;
    c -= (uint32_t)k;
  }
  return s;
}

}  // namespace g3a
