// Learned-index form: the per-leaf training steps, host/device code shared
// by the build kernels (g3a_learned.hip) and the host-side tests.
#pragma once
#include "g3a_dev.h"

namespace g3a {

// Leaf set-up from its row range [a, b) of the sorted keys: first row, key
// range and knots at the key quantiles (rows a + (h + 1)(b - a)/(kLiHinge + 1)).
G3A_HD void li_leaf_init(const uint64_t *keys, uint32_t a, uint32_t b, LiLeaf &L) {
  const int64_t c = (int64_t)b - a;
  L.a = a;
  L.err = 0;
  L.kmin = c ? keys[a] : 0;
  const uint64_t span = c ? keys[a + c - 1] - L.kmin : 0;
  L.inv_span = span ? 1.f / (float)span : 0.f;
  L.c0 = L.c1 = L.pad = 0.f;
  for (int h = 0; h < kLiHinge; ++h) {
    const int64_t q = a + (h + 1) * c / (kLiHinge + 1);
    L.t[h] = c ? li_leaf_x(L, keys[q < a + c ? q : a + c - 1]) : 0.f;
    L.w[h] = 0.f;
  }
}

// Segment of row r in a leaf of rows [a, a + c): the number of knot rows <= r.
// On segment g the hidden units h < g are active (x >= t_h, keys being
// sorted) and the others are 0.
G3A_HD int li_segment(int64_t a, int64_t c, int64_t r) {
  int g = 0;
  for (int h = 0; h < kLiHinge; ++h) g += (a + (h + 1) * c / (kLiHinge + 1) <= r) ? 1 : 0;
  return g;
}

// Output layer of one leaf: least squares of y = row - a over the leaf's
// distinct keys, from the moments m[5 g + (0..4)] = (count, sum x, sum x^2,
// sum y, sum x y) of each segment g.  On a segment every feature (1, x,
// x - t_h or 0) is affine in x, so the normal equations follow from the
// moments.  FP64 Gaussian elimination with partial pivoting and a small
// ridge; the weights are stored as FP32.
G3A_HD void li_solve_leaf(LiLeaf &L, const double *m) {
  constexpr int D = kLiHinge + 2;
  double A[D][D + 1];
  for (int a = 0; a < D; ++a)
    for (int b = 0; b <= D; ++b) A[a][b] = 0.0;
  for (int g = 0; g <= kLiHinge; ++g) {
    const double *mg = m + 5 * g;
    if (mg[0] == 0.0) continue;
    double al[D], be[D];
    al[0] = 1.0;
    be[0] = 0.0;
    al[1] = 0.0;
    be[1] = 1.0;
    for (int h = 0; h < kLiHinge; ++h) {
      const bool on = g >= h + 1;
      al[2 + h] = on ? -(double)L.t[h] : 0.0;
      be[2 + h] = on ? 1.0 : 0.0;
    }
    for (int a = 0; a < D; ++a) {
      for (int b = 0; b < D; ++b)
        A[a][b] += al[a] * al[b] * mg[0] + (al[a] * be[b] + al[b] * be[a]) * mg[1] + be[a] * be[b] * mg[2];
      A[a][D] += al[a] * mg[3] + be[a] * mg[4];
    }
  }
  double dmax = 0.0;
  for (int a = 0; a < D; ++a) dmax = A[a][a] > dmax ? A[a][a] : dmax;
  for (int a = 0; a < D; ++a) A[a][a] += 1e-9 * dmax + 1e-12;
  for (int c = 0; c < D; ++c) {
    int piv = c;
    for (int a = c + 1; a < D; ++a) {
      const double u = A[a][c] < 0 ? -A[a][c] : A[a][c], v = A[piv][c] < 0 ? -A[piv][c] : A[piv][c];
      piv = u > v ? a : piv;
    }
    if (piv != c)
      for (int b = 0; b <= D; ++b) {
        const double tmp = A[c][b];
        A[c][b] = A[piv][b];
        A[piv][b] = tmp;
      }
    const double d = A[c][c];
    if (d == 0.0) continue;
    for (int a = c + 1; a < D; ++a) {
      const double f = A[a][c] / d;
      for (int b = c; b <= D; ++b) A[a][b] -= f * A[c][b];
    }
  }
  double sol[D];
  for (int c = D - 1; c >= 0; --c) {
    double v = A[c][D];
    for (int b = c + 1; b < D; ++b) v -= A[c][b] * sol[b];
    sol[c] = A[c][c] != 0.0 ? v / A[c][c] : 0.0;
  }
  L.c0 = (float)sol[0];
  L.c1 = (float)sol[1];
  for (int h = 0; h < kLiHinge; ++h) L.w[h] = (float)sol[2 + h];
}

}  // namespace g3a
