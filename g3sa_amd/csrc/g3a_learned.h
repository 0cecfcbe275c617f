// Learned-index form: device buffers and the build entry (g3a_learned.hip).
#pragma once
#include "g3a_dev.h"

#include <hip/hip_runtime.h>

namespace g3a {

struct LearnedBufs {
  uint32_t *saf;     // n + 1 rows
  uint64_t *keys;    // n + 1 rows
  uint32_t *root;    // 2^kLiRootBits + 1
  LiLeaf *leaf;      // kLiLeaves
  LiKnot *knot;      // the knot pool (sized from learned_build_begin)
  uint32_t *leaf_a;  // kLiLeaves + 1 (scratch)
  uint32_t *shortt;  // 2 * kLiShort
  unsigned long long *ctr;  // 8 (scratch)
  int64_t n_samples;        // SA samples in the index
};

struct LearnedBuildStats {
  int64_t unfilled_rows, order_violations, bad_walks, short_rows;
  int64_t distinct_keys, err_sum, max_err, keys_over_window, leaves_over_window;
  int64_t knots;      // knots in use over all leaves
  int32_t max_knots;  // most knots in one leaf
  uint32_t leaf_w;
};

// Builds the full SA, the keys and the model from the narrow FM-index in x
// on stream s, in two calls: begin builds everything but the knots and
// returns the knot pool's size; end (b.knot holding that many knots) fits
// the leaves and measures their errors.  Both synchronise s on return.  0 on
// success; 1: the text has 2^32 rows or more, or no narrow layout; 2: a HIP
// error; 3: a consistency check failed (rows not covered by the SA walks,
// keys out of order).
int learned_build_begin(DevIndex x, const LearnedBufs &b, hipStream_t s, int64_t *total_knots);
int learned_build_end(DevIndex x, const LearnedBufs &b, hipStream_t s, LearnedBuildStats &st);

}  // namespace g3a
