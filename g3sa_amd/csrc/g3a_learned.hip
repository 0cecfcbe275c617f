// Learned-index form: construction and training kernels.
//
// From the FM-index already in HBM this builds, on the device:
//   1. the full suffix array: one walk per SA sample, backwards through the
//      text by LF steps until the next sampled row, writing every row it
//      passes (each row is written once; n LF steps in all);
//   2. the key of every row: the first 32 symbols of its suffix, 2 bits
//      each, zero past the text end (keys are non-decreasing in row order);
//   3. the root: the first row of every 6-symbol prefix;
//   4. the leaves: first rows (from the root's routing), then each leaf's
//      network trained by an error-bounded fit over its distinct keys
//      (li_fit_pass: greedy, a knot where a line can no longer stay within
//      250 rows of every key; a counting pass sizes the knot pool, a second
//      pass writes it), and the largest prediction error over the distinct
//      keys, measured with the lookup's own evaluation.
// The model is read by li_range / li_lower_bound (g3a_dev.h), whose window
// search is exact; the error bounds only size the windows.
#include "g3a_dev.h"
#include "g3a_learned.h"
#include "g3a_learned_fit.h"

#include <hip/hip_runtime.h>

#include <vector>

namespace g3a {

namespace {

constexpr int kChunk = 16;  // rows per thread in the row-range passes

__global__ void __launch_bounds__(256) k_li_fill(uint32_t *p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// 1. full SA: walk from sampled row j << sa_shift (whose value is known)
// through the rows of positions p - 1, p - 2, ... until a sampled row or
// position 0 (the '$' BWT row).
__global__ void __launch_bounds__(256) k_li_full_sa(DevIndex x, uint32_t *saf, int64_t n_samples,
                                                    unsigned long long *ctr) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_samples) return;
  if (!li_walk_sample(x, j, saf)) atomicAdd(ctr + 7, 1ull);
}

// 2. keys, rows never written by the walks, and the short suffixes
__global__ void __launch_bounds__(256) k_li_keys(DevIndex x, const uint32_t *saf, uint64_t *keys, uint32_t *shortt,
                                                 unsigned long long *ctr) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > x.n) return;
  const uint32_t p = saf[r];
  if (p == 0xffffffffu || p > (uint32_t)x.n) {
    atomicAdd(ctr + 0, 1ull);
    keys[r] = 0;
    return;
  }
  keys[r] = text_word(x, p);
  if (x.n - p < kLiShort) {
    const unsigned long long t = atomicAdd(ctr + 6, 1ull);
    if (t < (unsigned long long)kLiShort) {
      shortt[2 * t] = (uint32_t)r;
      shortt[2 * t + 1] = (uint32_t)(x.n - p);
    }
  }
}

// order check (keys must be non-decreasing) and distinct-key count: a
// grid-stride loop with per-thread counts, one atomic per wavefront at the
// end (an atomic per wavefront per 64 rows serialises on the two counters)
__global__ void __launch_bounds__(256) k_li_order(const uint64_t *keys, int64_t N, unsigned long long *ctr) {
  unsigned long long nf = 0, nb = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[r], kp = r > 0 ? keys[r - 1] : 0;
    nf += (r == 0 || kp != k) ? 1 : 0;
    nb += (r > 0 && kp > k) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    nf += __shfl_xor(nf, o);
    nb += __shfl_xor(nb, o);
  }
  if ((threadIdx.x & 63) == 0) {
    if (nf) atomicAdd(ctr + 2, nf);
    if (nb) atomicAdd(ctr + 1, nb);
  }
}

// 3. root: li_root[s] = first row whose key's 6-symbol prefix is >= s;
// thread r writes the segments that start at row r (r = N: those after
// the last key)
__global__ void __launch_bounds__(256) k_li_root(const uint64_t *keys, int64_t N, uint32_t *root) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > N) return;
  const int64_t s1 = r == N ? (int64_t)1 << kLiRootBits : (int64_t)(keys[r] >> (64 - kLiRootBits));
  const int64_t s0 = r == 0 ? -1 : (int64_t)(keys[r - 1] >> (64 - kLiRootBits));
  for (int64_t s = s0 + 1; s <= s1; ++s) root[s] = (uint32_t)r;
}

// leaf first rows from the routing (monotone in the key)
__global__ void __launch_bounds__(256) k_li_bounds(DevIndex x, int64_t N, uint32_t *leaf_a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > N) return;
  const int64_t j1 = r == N ? kLiLeaves : (int64_t)li_route(x, x.keys[r]);
  const int64_t j0 = r == 0 ? -1 : (int64_t)li_route(x, x.keys[r - 1]);
  for (int64_t j = j0 + 1; j <= j1; ++j) leaf_a[j] = (uint32_t)r;
}

// leaf training, pass 1: knots per leaf (thread per leaf over its rows)
__global__ void __launch_bounds__(64) k_li_count(const uint64_t *keys, const uint32_t *leaf_a, LiLeaf *leaf) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kLiLeaves) return;
  LiLeaf L;
  L.a = leaf_a[j];
  L.err = 0;
  L.k0 = 0;
  L.nk = li_fit_pass(keys, leaf_a[j], leaf_a[j + 1], kLiFitEps, nullptr);
  leaf[j] = L;
}

// pool offsets of the leaves' knots (one thread: 4096 entries) and the total
__global__ void k_li_offsets(LiLeaf *leaf, unsigned long long *total) {
  if (blockIdx.x || threadIdx.x) return;
  unsigned long long t = 0;
  for (int j = 0; j < kLiLeaves; ++j) {
    leaf[j].k0 = (uint32_t)t;
    t += leaf[j].nk;
  }
  *total = t;
}

// pass 2: the knots, written at the leaf's pool offset
__global__ void __launch_bounds__(64) k_li_fit(const uint64_t *keys, const uint32_t *leaf_a, const LiLeaf *leaf,
                                               LiKnot *pool) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kLiLeaves) return;
  li_fit_pass(keys, leaf_a[j], leaf_a[j + 1], kLiFitEps, pool + leaf[j].k0);
}

// largest prediction error per leaf over its distinct keys (the same
// evaluation as the lookup), their sum and the keys a window would miss
__global__ void __launch_bounds__(256) k_li_err(DevIndex x, LiLeaf *leaf, int64_t N, unsigned long long *ctr) {
  const int64_t r0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kChunk;
  if (r0 >= N) return;
  const int64_t r1 = r0 + kChunk < N ? r0 + kChunk : N;
  int cur = -1;
  uint32_t emax = 0;
  unsigned long long esum = 0, over = 0;
  for (int64_t r = r0; r < r1; ++r) {
    const uint64_t K = x.keys[r];
    if (r > 0 && x.keys[r - 1] == K) continue;
    const int j = (int)li_route(x, K);
    const int64_t p = li_predict(x.li_knot, x.li_leaf[j], K);
    const int64_t d = p > r ? p - r : r - p;
    const uint32_t e = d < 0x7fffffff ? (uint32_t)d : 0x7fffffffu;
    if (j != cur) {
      if (cur >= 0) atomicMax(&leaf[cur].err, emax);
      cur = j;
      emax = 0;
    }
    emax = e > emax ? e : emax;
    esum += e;
    over += (e + 2 > (uint32_t)kLiWindowCap) ? 1 : 0;
  }
  if (cur >= 0) atomicMax(&leaf[cur].err, emax);
  if (esum) atomicAdd(ctr + 4, esum);
  if (over) atomicAdd(ctr + 5, over);
}

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int learned_build_begin(DevIndex x, const LearnedBufs &b, hipStream_t s, int64_t *total_knots) {
  const int64_t N = x.n + 1;
  if (N >= ((int64_t)1 << 32) || !x.occn) return 1;
  unsigned long long *ctr = b.ctr;
  if (hipMemsetAsync(ctr, 0, 8 * sizeof(unsigned long long), s) != hipSuccess) return 2;
  hipLaunchKernelGGL(k_li_fill, dim3(blocks(N, 256)), dim3(256), 0, s, b.saf, N, 0xffffffffu);
  hipLaunchKernelGGL(k_li_fill, dim3(1), dim3(256), 0, s, b.shortt, (int64_t)(2 * kLiShort), 0xffffffffu);
  const int64_t n_samples = b.n_samples;
  hipLaunchKernelGGL(k_li_full_sa, dim3(blocks(n_samples, 256)), dim3(256), 0, s, x, b.saf, n_samples, ctr);
  hipLaunchKernelGGL(k_li_keys, dim3(blocks(N, 256)), dim3(256), 0, s, x, (const uint32_t *)b.saf, b.keys, b.shortt,
                     ctr);
  hipLaunchKernelGGL(k_li_order, dim3(4096), dim3(256), 0, s, (const uint64_t *)b.keys, N, ctr);
  hipLaunchKernelGGL(k_li_root, dim3(blocks(N + 1, 256)), dim3(256), 0, s, (const uint64_t *)b.keys, N, b.root);
  x.saf = b.saf;
  x.keys = b.keys;
  x.li_root = b.root;
  x.li_leaf_w = (uint32_t)((N + kLiLeaves - 1) / kLiLeaves);
  hipLaunchKernelGGL(k_li_bounds, dim3(blocks(N + 1, 256)), dim3(256), 0, s, x, N, b.leaf_a);
  hipLaunchKernelGGL(k_li_count, dim3(blocks(kLiLeaves, 64)), dim3(64), 0, s, (const uint64_t *)b.keys,
                     (const uint32_t *)b.leaf_a, b.leaf);
  hipLaunchKernelGGL(k_li_offsets, dim3(1), dim3(64), 0, s, b.leaf, ctr + 3);
  if (hipGetLastError() != hipSuccess) return 2;
  unsigned long long t = 0;
  if (hipMemcpyAsync(&t, ctr + 3, sizeof t, hipMemcpyDeviceToHost, s) != hipSuccess) return 2;
  if (hipStreamSynchronize(s) != hipSuccess) return 2;
  *total_knots = (int64_t)t;
  return 0;
}

int learned_build_end(DevIndex x, const LearnedBufs &b, hipStream_t s, LearnedBuildStats &st) {
  const int64_t N = x.n + 1;
  unsigned long long *ctr = b.ctr;
  x.saf = b.saf;
  x.keys = b.keys;
  x.li_root = b.root;
  x.li_leaf = b.leaf;
  x.li_knot = b.knot;
  x.li_short = b.shortt;
  x.li_leaf_w = (uint32_t)((N + kLiLeaves - 1) / kLiLeaves);
  hipLaunchKernelGGL(k_li_fit, dim3(blocks(kLiLeaves, 64)), dim3(64), 0, s, (const uint64_t *)b.keys,
                     (const uint32_t *)b.leaf_a, (const LiLeaf *)b.leaf, b.knot);
  hipLaunchKernelGGL(k_li_err, dim3(blocks(blocks(N, kChunk), 256)), dim3(256), 0, s, x, b.leaf, N, ctr);
  if (hipGetLastError() != hipSuccess) return 2;
  unsigned long long c[8];
  if (hipMemcpyAsync(c, ctr, sizeof c, hipMemcpyDeviceToHost, s) != hipSuccess) return 2;
  std::vector<LiLeaf> leaves(kLiLeaves);
  if (hipMemcpyAsync(leaves.data(), b.leaf, sizeof(LiLeaf) * kLiLeaves, hipMemcpyDeviceToHost, s) != hipSuccess)
    return 2;
  if (hipStreamSynchronize(s) != hipSuccess) return 2;
  st.unfilled_rows = (int64_t)c[0];
  st.order_violations = (int64_t)c[1];
  st.distinct_keys = (int64_t)c[2];
  st.knots = (int64_t)c[3];
  st.err_sum = (int64_t)c[4];
  st.keys_over_window = (int64_t)c[5];
  st.short_rows = (int64_t)c[6];
  st.bad_walks = (int64_t)c[7];
  st.max_err = 0;
  st.leaves_over_window = 0;
  st.max_knots = 0;
  for (const LiLeaf &L : leaves) {
    st.max_err = L.err > (uint32_t)st.max_err ? (int64_t)L.err : st.max_err;
    st.leaves_over_window += L.err + 2 > (uint32_t)kLiWindowCap ? 1 : 0;
    st.max_knots = (int32_t)L.nk > st.max_knots ? (int32_t)L.nk : st.max_knots;
  }
  st.leaf_w = x.li_leaf_w;
  const int64_t want_short = N < kLiShort ? N : kLiShort;
  if (st.unfilled_rows || st.order_violations || st.bad_walks || st.short_rows != want_short) return 3;
  return 0;
}

}  // namespace g3a
