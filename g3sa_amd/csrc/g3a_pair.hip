// Paired-end primitives on the device: the insert-size distribution of a
// batch of read pairs and the choice of the best pair of hits.  Reads 2p and
// 2p+1 of a batch are the two ends of pair p; each end's hits are its
// alignment list (score-sorted, best first) as the pipeline leaves it.
//
// The reference carries these as device ports it never launches:
//   mem_infer_dir  bwa/src/aux.cu:411-418
//   cal_sub        bwa/src/aux.cu:420-431
//   mem_pestat     bwa/src/aux.cu:435-500 (under #if 0)
//   mem_pair       bwa/src/aux.cu:1000-1070 (commented out)
//   proper-pair    bwa/src/aux.cu:1288-1295 (commented out mem_sam_pe)
//   hash_64        bwa/src/aux.cu:703-714
// Floating-point expressions keep the reference's evaluation order and are
// compiled without contraction, so that the statistics come out bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "g3a_kernels.h"
#include "g3a_pair.h"

#pragma clang fp contract(off)

namespace g3a {

namespace {

// bwa/src/aux.cuh:404-409
constexpr double kMinRatio = 0.8, kMinDirRatio = 0.05, kOutlierBound = 2.0, kMappingBound = 3.0,
                 kMaxStddev = 4.0;
constexpr int kMinDirCnt = 10;

// Orientation of two hits (0 FF, 1 FR, 2 RF, 3 RR) and their distance on the
// strand of the first (bwa/src/aux.cu:411-418).
__device__ __forceinline__ int infer_dir(int64_t l_pac, int64_t b1, int64_t b2, int64_t *dist) {
  const int r1 = b1 >= l_pac, r2 = b2 >= l_pac;
  const int64_t p2 = r1 == r2 ? b2 : (l_pac << 1) - 1 - b2;
  *dist = p2 > b1 ? p2 - b1 : b1 - p2;
  return (r1 == r2 ? 0 : 1) ^ (p2 > b1 ? 0 : 3);
}

// Score of the best hit that overlaps the top hit on the read by at least
// mask_level of the shorter, else min_seed_len * a (bwa/src/aux.cu:420-431).
__device__ int cal_sub(const PairArgs &a, const Aln *r, int64_t n) {
  int64_t j;
  for (j = 1; j < n; ++j) {
    const int b_max = r[j].qb > r[0].qb ? r[j].qb : r[0].qb;
    const int e_min = r[j].qe < r[0].qe ? r[j].qe : r[0].qe;
    if (e_min > b_max) {
      const int lj = r[j].qe - r[j].qb, l0 = r[0].qe - r[0].qb;
      const int min_l = lj < l0 ? lj : l0;
      if ((float)(e_min - b_max) >= (float)min_l * a.mask_level) break;
    }
  }
  return j < n ? r[j].score : a.min_seed_len * a.a;
}

// Thomas Wang's 64-bit integer mix (bwa/src/aux.cu:703-714).
__device__ __forceinline__ uint64_t hash_64(uint64_t key) {
  key += ~(key << 32);
  key ^= key >> 22;
  key += ~(key << 13);
  key ^= key >> 8;
  key += key << 3;
  key ^= key >> 15;
  key += ~(key << 27);
  key ^= key >> 31;
  return key;
}

__global__ void __launch_bounds__(256) k_pe_zero(uint32_t *h, int64_t n) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) h[i] = 0;
}

// One thread per pair: the insert size of a pair whose ends both have a
// unique top hit on one contig, counted in its orientation's histogram
// (bwa/src/aux.cu:441-459).  The reference collects the sizes in per-
// orientation lists and sorts them; the histogram holds the same multiset.
// Sizes cluster on a few hundred words, so each block counts into one of
// `copies` histograms (atomics on one word serialise in L2), which k_pe_fold
// adds into the first.
__global__ void __launch_bounds__(256) k_pe_isize(PairArgs a) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n_pairs) return;
  const int64_t b0 = a.aln_off[2 * p], b1 = a.aln_off[2 * p + 1], e1 = a.aln_off[2 * p + 2];
  if (b1 == b0 || e1 == b1) return;
  const Aln *r0 = a.alns + b0, *r1 = a.alns + b1;
  if (cal_sub(a, r0, b1 - b0) > kMinRatio * r0->score) return;
  if (cal_sub(a, r1, e1 - b1) > kMinRatio * r1->score) return;
  if (r0->rid != r1->rid) return;
  int64_t is;
  const int dir = infer_dir(a.l_pac, r0->rb, r1->rb, &is);
  if (is && is <= a.max_ins)
    atomicAdd(a.hist + ((int64_t)(blockIdx.x % a.copies) * 4 + dir) * (a.max_ins + 1) + is, 1u);
}

__global__ void __launch_bounds__(256) k_pe_fold(PairArgs a) {
  const int64_t nb = 4 * ((int64_t)a.max_ins + 1);
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < nb; i += (int64_t)gridDim.x * 256) {
    uint32_t s = a.hist[i];
    for (int c = 1; c < a.copies; ++c) s += a.hist[c * nb + i];
    a.hist[i] = s;
  }
}

// Inclusive prefix sum over a wavefront.
__device__ __forceinline__ int64_t wave_scan(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int64_t u = __shfl_up(v, d);
    if (lane >= d) v += u;
  }
  return v;
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
#pragma unroll
  for (int d = 32; d; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// One wavefront per orientation (bwa/src/aux.cu:461-491): quartiles of the
// sizes, the mean and standard deviation of those within the outlier bounds,
// and the mapping bounds.  The mean's running sum is exact (integers below
// 2^53), so it is summed here as an integer; the variance's running sum
// rounds at every term, in the reference's order (ascending size): the c
// equal terms of one size are added by repeat_add, which gives the result of
// c single rounded adds.  Then the orientations with too few pairs against
// the most frequent one fail (bwa/src/aux.cu:492-498).
__global__ void __launch_bounds__(256) k_pe_stat(PairArgs a) {
  __shared__ int64_t cnt[4];
  const int d = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t M = a.max_ins;
  const uint32_t *h = a.hist + (int64_t)d * (M + 1);
  PeStat st{};
  int64_t n = 0;
  for (int64_t v = lane; v <= M; v += 64) n += h[v];
  n = wave_sum(n);
  if (lane == 0) cnt[d] = n;
  if (n < kMinDirCnt) {
    st.failed = 1;
  } else {
    const int64_t t[3] = {(int64_t)(.25 * n + .499), (int64_t)(.50 * n + .499), (int64_t)(.75 * n + .499)};
    int64_t pv[3] = {0, 0, 0};
    int64_t run = 0;
    for (int64_t base = 0; base <= M; base += 64) {
      const int64_t v = base + lane;
      const int64_t c = v <= M ? h[v] : 0;
      const int64_t inc = wave_scan(c) + run;
      const int64_t exc = inc - c;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const uint64_t hit = __ballot(exc <= t[k] && t[k] < inc);
        if (hit) pv[k] = base + __ffsll((unsigned long long)hit) - 1;
      }
      run = __shfl(inc, 63);
    }
    const int p25 = (int)pv[0], p75 = (int)pv[2];
    int low = (int)(p25 - kOutlierBound * (p75 - p25) + .499);
    if (low < 1) low = 1;
    const int high = (int)(p75 + kOutlierBound * (p75 - p25) + .499);
    int64_t sum = 0, x = 0;
    for (int64_t v = low + lane; v <= high && v <= M; v += 64) {
      const int64_t c = h[v];
      sum += c * v;
      x += c;
    }
    sum = wave_sum(sum);
    x = wave_sum(x);
    double avg = (double)sum;
    avg /= (double)x;
    // every lane runs the same sum; the wavefront loads 64 counts at a time
    const int64_t hi = high < M ? high : M;
    double sd = 0;
    for (int64_t base = low; base <= hi; base += 64) {
      const uint32_t c = base + lane <= hi ? h[base + lane] : 0;
      const int nj = (int)(hi - base + 1 < 64 ? hi - base + 1 : 64);
      for (int j = 0; j < nj; ++j) {
        const uint32_t cj = __shfl(c, j);
        if (cj) {
          const double dv = (double)(base + j) - avg;
          sd = repeat_add(sd, dv * dv, cj);
        }
      }
    }
    sd = sqrt(sd / (double)x);
    st.avg = avg;
    st.std = sd;
    st.low = (int)(p25 - kMappingBound * (p75 - p25) + .499);
    st.high = (int)(p75 + kMappingBound * (p75 - p25) + .499);
    if (st.low > avg - kMaxStddev * sd) st.low = (int)(avg - kMaxStddev * sd + .499);
    if (st.high < avg + kMaxStddev * sd) st.high = (int)(avg + kMaxStddev * sd + .499);
    if (st.low < 1) st.low = 1;
  }
  __syncthreads();
  if (lane == 0) {
    int64_t mx = 0;
    for (int k = 0; k < 4; ++k) mx = mx > cnt[k] ? mx : cnt[k];
    if (st.failed == 0 && (double)n < mx * kMinDirRatio) st.failed = 1;
    a.stat[d] = st;
  }
}

// Pair score of hits k (earlier) and i of the position-sorted hit list, as
// the reference rounds it (bwa/src/aux.cu:1041-1043).
__device__ __forceinline__ int pair_q(const PairArgs &a, const PeStat &ps, uint64_t yi, uint64_t yk, int64_t dist) {
  const double ns = ((double)dist - ps.avg) / ps.std;
  const double sc = (double)((yi >> 32) + (yk >> 32));
  const double qd = sc + .721 * log(2. * erfc(fabs(ns) * 0.70710678118654752440)) * a.a + .499;
  return qd > 0 ? (int)qd : 0;  // negative, -inf and NaN all end as 0 in the reference
}

// Visit every consistent pair (k, i) of the sorted hits in the reference's
// order (bwa/src/aux.cu:1021-1055), calling f(q, y) with y = k << 32 | i.
template <class F>
__device__ void pair_walk(const PairArgs &a, const uint64_t *v, int nv, F &&f) {
  int last[4] = {-1, -1, -1, -1};
  for (int i = 0; i < nv; ++i) {
    const uint64_t xi = v[2 * i], yi = v[2 * i + 1];
    for (int r = 0; r < 2; ++r) {
      const int dir = r << 1 | (int)(yi >> 1 & 1);
      const PeStat &ps = a.pes[dir];
      if (ps.failed) continue;
      const int which = r << 1 | (int)((yi & 1) ^ 1);
      if (last[which] < 0) continue;
      for (int k = last[which]; k >= 0; --k) {
        const uint64_t yk = v[2 * k + 1];
        if ((int)(yk & 3) != which) continue;
        const int64_t dist = (int64_t)xi - (int64_t)v[2 * k];
        if (dist > ps.high) break;
        if (dist < ps.low) continue;
        f(pair_q(a, ps, yi, yk, dist), (uint64_t)k << 32 | (uint32_t)i);
      }
    }
    last[yi & 3] = i;
  }
}

// One thread per pair: the best pair of hits of the two ends (mem_pair,
// bwa/src/aux.cu:1000-1070) and the proper-pair flag of the two top hits
// (bwa/src/aux.cu:1288-1295).  Output per pair: score of the best pair (0 if
// none), second-best score, number of near-best pairs, hit index of read 1
// and of read 2 (-1 if none), and 2 if the top hits form a proper pair.
// The hit keys of a pair live in the workspace slice of its alignments.
__global__ void __launch_bounds__(256) k_pe_pair(PairArgs a) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n_pairs) return;
  const int64_t b[3] = {a.aln_off[2 * p], a.aln_off[2 * p + 1], a.aln_off[2 * p + 2]};
  int32_t *o = a.out + 6 * p;
  int best_sc = 0, sub = 0, n_sub = 0, z[2] = {-1, -1}, proper = 0;
  const int64_t n0 = b[1] - b[0], n1 = b[2] - b[1];
  if (n0 > 0 && n1 > 0) {
    const Aln &t0 = a.alns[b[0]], &t1 = a.alns[b[1]];
    if (t0.rid == t1.rid && t0.rid >= 0) {
      int64_t dist;
      const int d = infer_dir(a.l_pac, t0.rb, t1.rb, &dist);
      if (!a.pes[d].failed && dist >= a.pes[d].low && dist <= a.pes[d].high) proper = 2;
    }
    // hit keys: x = contig << 32 | forward position in the contig,
    // y = score << 32 | hit index << 2 | strand << 1 | end
    uint64_t *v = a.v + 2 * b[0];
    const int nv = (int)(b[2] - b[0]);
    for (int e = 0, m = 0; e < 2; ++e)
      for (int64_t i = 0; i < b[e + 1] - b[e]; ++i, ++m) {
        const Aln &h = a.alns[b[e] + i];
        const int64_t fwd = h.rb < a.l_pac ? h.rb : (a.l_pac << 1) - 1 - h.rb;
        const uint64_t x = (uint64_t)h.rid << 32 | (uint64_t)(fwd - a.ctg_off[h.rid]);
        const uint64_t y = (uint64_t)h.score << 32 | (uint64_t)(int64_t)((int)i << 2 | (h.rb >= a.l_pac) << 1 | e);
        int j = m;  // insertion into the sorted prefix, ascending by (x, y)
        while (j > 0 && (v[2 * j - 2] > x || (v[2 * j - 2] == x && v[2 * j - 1] > y))) {
          v[2 * j] = v[2 * j - 2];
          v[2 * j + 1] = v[2 * j - 1];
          --j;
        }
        v[2 * j] = x;
        v[2 * j + 1] = y;
      }
    // the reference's pair keys: q << 32 | low half of hash(y ^ id << 8),
    // sorted ascending; the last is the best, the one before it the second
    const int id = (int)(a.id0 + p);
    const uint64_t idk = (uint64_t)(int64_t)(int32_t)((uint32_t)id << 8);
    uint64_t bx = 0, by = 0, sx = 0, sy = 0;
    int64_t nu = 0;
    pair_walk(a, v, nv, [&](int q, uint64_t y) {
      const uint64_t x = (uint64_t)q << 32 | (hash_64(y ^ idk) & 0xffffffffu);
      if (nu == 0 || x > bx || (x == bx && y > by)) {
        sx = bx, sy = by;
        bx = x, by = y;
      } else if (nu == 1 || x > sx || (x == sx && y > sy)) {
        sx = x, sy = y;
      }
      ++nu;
    });
    if (nu) {
      int tmp = a.a + a.b;
      tmp = tmp > a.o_del + a.e_del ? tmp : a.o_del + a.e_del;
      tmp = tmp > a.o_ins + a.e_ins ? tmp : a.o_ins + a.e_ins;
      const int i = (int)(by >> 32), k = (int)(by & 0xffffffffu);
      z[v[2 * i + 1] & 1] = (int)((v[2 * i + 1] & 0xffffffffu) >> 2);
      z[v[2 * k + 1] & 1] = (int)((v[2 * k + 1] & 0xffffffffu) >> 2);
      best_sc = (int)(bx >> 32);
      sub = nu > 1 ? (int)(sx >> 32) : 0;
      pair_walk(a, v, nv, [&](int q, uint64_t y) {
        if (y != by && sub - q <= tmp) ++n_sub;
      });
    }
  }
  o[0] = best_sc;
  o[1] = sub;
  o[2] = n_sub;
  o[3] = z[0];
  o[4] = z[1];
  o[5] = proper;
}

}  // namespace

void launch_pe_stat(const PairArgs &a, hipStream_t s) {
  const int64_t nb = 4 * ((int64_t)a.max_ins + 1), nh = nb * a.copies;
  hipLaunchKernelGGL(k_pe_zero, dim3((unsigned)std::min<int64_t>((nh + 255) / 256, 4096)), dim3(256), 0, s, a.hist, nh);
  if (a.n_pairs)
    hipLaunchKernelGGL(k_pe_isize, dim3((unsigned)((a.n_pairs + 255) / 256)), dim3(256), 0, s, a);
  if (a.copies > 1)
    hipLaunchKernelGGL(k_pe_fold, dim3((unsigned)std::min<int64_t>((nb + 255) / 256, 1024)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_pe_stat, dim3(1), dim3(256), 0, s, a);
}

void launch_pe_pair(const PairArgs &a, hipStream_t s) {
  if (a.n_pairs)
    hipLaunchKernelGGL(k_pe_pair, dim3((unsigned)((a.n_pairs + 255) / 256)), dim3(256), 0, s, a);
}

}  // namespace g3a
