// Learned-index form: training of one leaf, host/device code shared by the
// build kernels (g3a_learned.hip) and the host-side tests.
#pragma once
#include "g3a_dev.h"

namespace g3a {

constexpr double kLiFitEps = 250.0;  // error the fit aims for (window half-width is err + 2 <= 256)

// The leaf's network fitted to its distinct keys (rows [a, b), the target
// y = row - a) by a greedy error-bounded pass: from the current knot
// (X0, Y0), the slopes whose line stays within eps of every key seen so far
// form an interval [lo, hi], which each key narrows.  When a key would empty
// it, the segment ends at the previous key with the middle slope, and the
// next segment starts from that knot, so the function stays continuous and
// every key lies within eps of it.  Key differences are taken in 64-bit
// integers and the cone in FP64; the knots are stored with FP32 outputs and
// slopes.  Writes the knots to out (when not null) and returns their number.
G3A_HD uint32_t li_fit_pass(const uint64_t *keys, uint32_t a, uint32_t b, double eps, LiKnot *out) {
  uint32_t nk = 0;
  uint64_t X0 = 0, last = 0;
  double Y0 = 0, lo = -1e300, hi = 1e300;
  bool have = false;
  for (uint32_t r = a; r < b; ++r) {
    if (r > a && keys[r] == keys[r - 1]) continue;
    const uint64_t K = keys[r];
    const double y = (double)(r - a);
    if (!have) {
      X0 = last = K;
      Y0 = y;
      have = true;
      continue;
    }
    double dx = (double)(K - X0);
    const double nlo = (y - eps - Y0) / dx, nhi = (y + eps - Y0) / dx;
    const double clo = nlo > lo ? nlo : lo, chi = nhi < hi ? nhi : hi;
    if (clo <= chi) {
      lo = clo;
      hi = chi;
      last = K;
      continue;
    }
    const double sl = 0.5 * (lo + hi);
    if (out) out[nk] = LiKnot{X0, (float)Y0, (float)sl};
    ++nk;
    Y0 += sl * (double)(last - X0);
    X0 = last;
    dx = (double)(K - X0);
    lo = (y - eps - Y0) / dx;
    hi = (y + eps - Y0) / dx;
    last = K;
  }
  if (have) {
    const double sl = lo > -1e299 && hi < 1e299 ? 0.5 * (lo + hi) : 0.0;
    if (out) out[nk] = LiKnot{X0, (float)Y0, (float)sl};
    ++nk;
  }
  return nk;
}

}  // namespace g3a
