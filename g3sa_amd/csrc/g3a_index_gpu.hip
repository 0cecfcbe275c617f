// Device construction of the FM-index files (.bwt .sa .counts .g3sa) on one
// MI355X, for the text fwd + revcomp + '$' of a reference.
//
// The reference ships no suffix sorter of its own that runs on its GPU: its
// index comes from bwa-mem2's host builder (bwa/src/FMI_search.cpp:232-420,
// bwt.c:62-84).  Here the suffix array is built by prefix doubling over
// radix sorts (rocPRIM) in HBM, then the BWT, the BWA block layout, the SA
// samples and the count tables are produced by device kernels and copied out
// once.  The files are byte-identical to the host SA-IS build
// (index_build.cpp), which the reference's E. coli fixtures pin.
//
// Prefix doubling, as run here:
//   round 0  every suffix i gets the key of its first 21 symbols (3 bits
//            each; '$' = 0, A..T = 1..4, zero past the text end) and the
//            (key, i) pairs are radix-sorted.  A suffix whose first 21
//            symbols hold '$' has a unique key, so it is placed for good.
//   group    a run of equal keys is a group; rank[i] = the SA row where
//            i's group starts.  Groups of one suffix are final.
//   round h  (h = 21, 42, 84, ...) only the suffixes of unfinished groups
//            take part.  Their key is (dense group number, rank[i + h]),
//            radix-sorted; a run of equal keys is a new group, whose first
//            row becomes the new rank of its members.  A suffix of an
//            unfinished group shares its first h symbols with another
//            suffix, so '$' is not among them and i + h <= n.
// The loop ends when every group is a single suffix.  Each round reads the
// ranks of the previous one (the keys are built before any rank changes).
#include "g3a_index.h"

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace g3a {

void (*device_build_progress)(const char *what, int level, int64_t n) = nullptr;

namespace {

#define DCK(call)                                                                   \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      err = std::string(#call ": ") + hipGetErrorString(e_);                        \
      return false;                                                                 \
    }                                                                               \
  } while (0)

// Device allocations of one build, released together.
struct Arena {
  std::vector<void *> ptrs;
  ~Arena() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  bool get(T *&p, size_t n, std::string &err) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
      err = std::string("device allocation of ") + std::to_string(n * sizeof(T)) + " bytes: " + hipGetErrorString(e);
      return false;
    }
    ptrs.push_back(q);
    p = (T *)q;
    return true;
  }
};

constexpr int kInitSyms = 21;  // symbols in the round-0 key (3 bits each)
constexpr int kBlock = 256;

inline unsigned grid_for(uint64_t n) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)(b < 65536 ? (b ? b : 1) : 65536);
}

inline int bits_for(uint64_t v) {  // bits needed to hold values 0..v
  int b = 1;
  while (b < 64 && (v >> b)) ++b;
  return b;
}

#define GRID_STRIDE(i, n) \
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

// text symbols: T[i] = fwd[i] + 1, T[2L-1-i] = 4 - fwd[i] (the complement
// strand read backwards), T[2L] = 0 ('$')
__global__ void k_text(const uint8_t *fwd, uint64_t L, uint8_t *T) {
  GRID_STRIDE(i, L) {
    const uint8_t b = fwd[i] & 3;
    T[i] = (uint8_t)(b + 1);
    T[2 * L - 1 - i] = (uint8_t)(4 - b);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) T[2 * L] = 0;
}

template <class I>
__global__ void k_init_keys(const uint8_t *T, uint64_t N, uint64_t *key, I *val) {
  GRID_STRIDE(i, N) {
    uint64_t k = 0;
#pragma unroll
    for (int j = 0; j < kInitSyms; ++j) k = k << 3 | (i + j < N ? T[i + j] : 0u);
    key[i] = k;
    val[i] = (I)i;
  }
}

// hv[r] = r where a run of equal keys starts, else 0 (a max-scan of it gives
// each row its group's first row)
template <class I>
__global__ void k_head_input(const uint64_t *key, const I *rows, uint64_t m, I *hv) {
  GRID_STRIDE(p, m) hv[p] = (p == 0 || key[p] != key[p - 1]) ? rows[p] : (I)0;
}

// rank[i] = group start row of every sorted suffix i = sa_vals[p]; keep[p]
// marks members of groups of two or more
template <class I>
__global__ void k_rank(const uint64_t *key, const I *vals, const I *head, uint64_t m, I *rank, uint8_t *keep) {
  GRID_STRIDE(p, m) {
    rank[vals[p]] = head[p];
    const bool first = p == 0 || key[p] != key[p - 1];
    const bool last = p + 1 == m || key[p + 1] != key[p];
    keep[p] = !(first && last);
  }
}

template <class I>
__global__ void k_group_flag(const I *uhead, uint64_t m, I *gf) {
  GRID_STRIDE(p, m) gf[p] = (p == 0 || uhead[p] != uhead[p - 1]) ? (I)1 : (I)0;
}

// key = (dense group number, rank of the suffix h symbols on), val = suffix
template <class I>
__global__ void k_round_keys(const I *urows, const I *gid, const I *SA, const I *rank, uint64_t m, uint64_t h,
                             uint64_t N, int rank_bits, uint64_t *key, I *val) {
  GRID_STRIDE(p, m) {
    const uint64_t i = SA[urows[p]];
    // i + h <= n < N for every suffix of an unfinished group (see above);
    // the guard only keeps a broken invariant from reading out of bounds
    const uint64_t r2 = i + h < N ? (uint64_t)rank[i + h] : 0;
    key[p] = ((uint64_t)(gid[p] - 1) << rank_bits) | r2;
    val[p] = (I)i;
  }
}

template <class I>
__global__ void k_scatter_sa(const I *urows, const I *val, uint64_t m, I *SA) {
  GRID_STRIDE(p, m) SA[urows[p]] = val[p];
}

template <class I>
__global__ void k_find_primary(const I *SA, uint64_t N, uint64_t *primary) {
  GRID_STRIDE(r, N) if (SA[r] == 0) *primary = r;
}

// BWT string of BWA: position k is row k (k < primary) or row k + 1
template <class I>
__global__ void k_bwt_bytes(const uint8_t *T, const I *SA, const uint64_t *primary, uint64_t n, uint8_t *bw) {
  const uint64_t pr = *primary;
  GRID_STRIDE(k, n) {
    const uint64_t row = k < pr ? k : k + 1;
    bw[k] = (uint8_t)(T[(uint64_t)SA[row] - 1] - 1);
  }
}

// per-base counts of each 128-position block of the BWT string
__global__ void k_block_counts(const uint8_t *bw, uint64_t n, uint64_t nb, uint64_t *bc) {
  GRID_STRIDE(b, nb) {
    uint64_t c[4] = {0, 0, 0, 0};
    const uint64_t e = (b + 1) * 128 < n ? (b + 1) * 128 : n;
    for (uint64_t k = b * 128; k < e; ++k) ++c[bw[k] & 3];
    for (int x = 0; x < 4; ++x) bc[x * nb + b] = c[x];
  }
}

// BWA .bwt body: per 128-position block, the four uint64 counts before it
// (as eight little-endian uint32 words) then up to eight words of 16
// symbols, first symbol in the highest bits; after the last block the
// totals (bwt.c:385-419 layout, as index_build.cpp writes it)
__global__ void k_bwt_words(const uint8_t *bw, const uint64_t *occ, uint64_t n, uint64_t nb, uint32_t *words) {
  GRID_STRIDE(b, nb) {
    uint32_t *o = words + b * 16;
    for (int x = 0; x < 4; ++x) {
      const uint64_t v = occ[x * nb + b];
      o[2 * x] = (uint32_t)v;
      o[2 * x + 1] = (uint32_t)(v >> 32);
    }
    for (int j = 0; j < 8; ++j) {
      const uint64_t k0 = b * 128 + (uint64_t)j * 16;
      if (k0 >= n) break;
      uint32_t w = 0;
      for (int t = 0; t < 16; ++t)
        if (k0 + t < n) w |= (uint32_t)(bw[k0 + t] & 3) << ((15 - t) << 1);
      o[8 + j] = w;
    }
  }
}

// base and dinucleotide counts of the text (bwa-mem2 .counts,
// FMI_search.cpp:248-259)
__global__ void k_text_counts(const uint8_t *T, uint64_t n, unsigned long long *cnt) {
  __shared__ unsigned long long h[20];
  if (threadIdx.x < 20) h[threadIdx.x] = 0;
  __syncthreads();
  GRID_STRIDE(i, n) {
    const int a = T[i] - 1;
    atomicAdd(&h[a], 1ULL);
    if (i + 1 < n) atomicAdd(&h[4 + a * 4 + (T[i + 1] - 1)], 1ULL);
  }
  __syncthreads();
  if (threadIdx.x < 20 && h[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], h[threadIdx.x]);
}

template <class I, class O>
__global__ void k_sample(const I *SA, uint64_t first, uint64_t step, uint64_t count, O *out) {
  GRID_STRIDE(i, count) out[i] = (O)SA[(first + i) * step];
}

void progress(const char *what, int level, int64_t n) {
  if (device_build_progress) device_build_progress(what, level, n);
}

// rows 0..m-1 in order (round 0's rows are the identity)
template <class I>
__global__ void k_head_input_iota(const uint64_t *key, uint64_t m, I *hv) {
  GRID_STRIDE(p, m) hv[p] = (p == 0 || key[p] != key[p - 1]) ? (I)p : (I)0;
}

template <class I>
bool build_device(const std::vector<uint8_t> &fwd, int dense_intv, FmFiles &out, std::string &err) {
  const uint64_t L = fwd.size(), n = 2 * L, N = n + 1;
  hipStream_t s = nullptr;
  Arena A;
  uint8_t *d_fwd, *T, *keep;
  I *SA, *rank, *valb[2], *urows[2], *uhead[2], *hv, *tI;
  uint64_t *keyb[2];
  size_t *d_count;
  if (!A.get(d_fwd, L, err) || !A.get(T, N + 32, err) || !A.get(SA, N, err) || !A.get(rank, N, err) ||
      !A.get(keyb[0], N, err) || !A.get(keyb[1], N, err) || !A.get(valb[0], N, err) || !A.get(valb[1], N, err) ||
      !A.get(urows[0], N, err) || !A.get(urows[1], N, err) || !A.get(uhead[0], N, err) ||
      !A.get(uhead[1], N, err) || !A.get(hv, N, err) || !A.get(tI, N, err) || !A.get(keep, N, err) ||
      !A.get(d_count, 1, err))
    return false;
  DCK(hipMemcpy(d_fwd, fwd.data(), L, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_text, dim3(grid_for(L)), dim3(kBlock), 0, s, d_fwd, L, T);
  hipLaunchKernelGGL(k_init_keys<I>, dim3(grid_for(N)), dim3(kBlock), 0, s, T, N, keyb[0], valb[0]);
  DCK(hipGetLastError());
  // temporary storage, sized for the largest call of each primitive
  size_t tmp_b = 0;
  {
    size_t b = 0;
    rocprim::double_buffer<uint64_t> kd(keyb[0], keyb[1]);
    rocprim::double_buffer<I> vd(valb[0], valb[1]);
    DCK(rocprim::radix_sort_pairs(nullptr, b, kd, vd, (size_t)N, 0, 64, s));
    tmp_b = b > tmp_b ? b : tmp_b;
    DCK(rocprim::inclusive_scan(nullptr, b, hv, tI, (size_t)N, rocprim::maximum<I>(), s));
    tmp_b = b > tmp_b ? b : tmp_b;
    DCK(rocprim::inclusive_scan(nullptr, b, hv, tI, (size_t)N, rocprim::plus<I>(), s));
    tmp_b = b > tmp_b ? b : tmp_b;
    DCK(rocprim::select(nullptr, b, rocprim::counting_iterator<I>(0), keep, urows[0], d_count, (size_t)N, s));
    tmp_b = b > tmp_b ? b : tmp_b;
    DCK(rocprim::select(nullptr, b, tI, keep, uhead[0], d_count, (size_t)N, s));
    tmp_b = b > tmp_b ? b : tmp_b;
    DCK(rocprim::exclusive_scan(nullptr, b, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint64_t)0, (size_t)N,
                                rocprim::plus<uint64_t>(), s));
    tmp_b = b > tmp_b ? b : tmp_b;
  }
  uint8_t *tmp;
  if (!A.get(tmp, tmp_b, err)) return false;
  const int rank_bits = bits_for(N - 1);
  // ---- round 0: every suffix by its first 21 symbols
  uint64_t *key;
  I *val;
  {
    rocprim::double_buffer<uint64_t> kd(keyb[0], keyb[1]);
    rocprim::double_buffer<I> vd(valb[0], valb[1]);
    size_t b = tmp_b;
    DCK(rocprim::radix_sort_pairs(tmp, b, kd, vd, (size_t)N, 0, 3 * kInitSyms, s));
    key = kd.current();
    val = vd.current();
  }
  DCK(hipMemcpyAsync(SA, val, N * sizeof(I), hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_head_input_iota<I>, dim3(grid_for(N)), dim3(kBlock), 0, s, key, N, hv);
  {
    size_t b = tmp_b;
    DCK(rocprim::inclusive_scan(tmp, b, hv, tI, (size_t)N, rocprim::maximum<I>(), s));
  }
  hipLaunchKernelGGL(k_rank<I>, dim3(grid_for(N)), dim3(kBlock), 0, s, key, val, tI, N, rank, keep);
  {
    size_t b = tmp_b;
    DCK(rocprim::select(tmp, b, rocprim::counting_iterator<I>(0), keep, urows[0], d_count, (size_t)N, s));
    b = tmp_b;
    DCK(rocprim::select(tmp, b, tI, keep, uhead[0], d_count, (size_t)N, s));
  }
  size_t m = 0;
  DCK(hipMemcpy(&m, d_count, sizeof m, hipMemcpyDeviceToHost));
  progress("suffixes sorted by 21 symbols; unsorted", 0, (int64_t)m);
  // ---- doubling rounds over the unfinished groups
  int u = 0, round = 1;
  for (uint64_t h = kInitSyms; m > 0; h *= 2, ++round) {
    hipLaunchKernelGGL(k_group_flag<I>, dim3(grid_for(m)), dim3(kBlock), 0, s, uhead[u], (uint64_t)m, hv);
    {
      size_t b = tmp_b;
      DCK(rocprim::inclusive_scan(tmp, b, hv, tI, m, rocprim::plus<I>(), s));
    }
    I groups = 0;
    DCK(hipMemcpy(&groups, tI + (m - 1), sizeof(I), hipMemcpyDeviceToHost));
    const int gbits = bits_for((uint64_t)groups - 1);
    if (gbits + rank_bits > 64) {
      err = "suffix sort: group and rank fields exceed a 64-bit key";
      return false;
    }
    hipLaunchKernelGGL(k_round_keys<I>, dim3(grid_for(m)), dim3(kBlock), 0, s, urows[u], tI, SA, rank, (uint64_t)m,
                       h, N, rank_bits, keyb[0], valb[0]);
    {
      rocprim::double_buffer<uint64_t> kd(keyb[0], keyb[1]);
      rocprim::double_buffer<I> vd(valb[0], valb[1]);
      size_t b = tmp_b;
      DCK(rocprim::radix_sort_pairs(tmp, b, kd, vd, m, 0, (unsigned)(gbits + rank_bits), s));
      key = kd.current();
      val = vd.current();
    }
    hipLaunchKernelGGL(k_scatter_sa<I>, dim3(grid_for(m)), dim3(kBlock), 0, s, urows[u], val, (uint64_t)m, SA);
    hipLaunchKernelGGL(k_head_input<I>, dim3(grid_for(m)), dim3(kBlock), 0, s, key, urows[u], (uint64_t)m, hv);
    {
      size_t b = tmp_b;
      DCK(rocprim::inclusive_scan(tmp, b, hv, tI, m, rocprim::maximum<I>(), s));
    }
    hipLaunchKernelGGL(k_rank<I>, dim3(grid_for(m)), dim3(kBlock), 0, s, key, val, tI, (uint64_t)m, rank, keep);
    {
      size_t b = tmp_b;
      DCK(rocprim::select(tmp, b, urows[u], keep, urows[u ^ 1], d_count, m, s));
      b = tmp_b;
      DCK(rocprim::select(tmp, b, tI, keep, uhead[u ^ 1], d_count, m, s));
    }
    DCK(hipMemcpy(&m, d_count, sizeof m, hipMemcpyDeviceToHost));
    u ^= 1;
    progress("doubling round; unsorted", round, (int64_t)m);
    if (round > 64) {
      err = "suffix sort did not converge";
      return false;
    }
  }
  DCK(hipGetLastError());
  // ---- BWT, BWA blocks, samples, counts
  uint64_t *d_primary, *bc, *occ;
  unsigned long long *d_cnt;
  const uint64_t nb = (n + 127) / 128;
  const uint64_t raw_words = (n + 15) / 16, n_occ = nb + 1, words = raw_words + n_occ * 8;
  uint32_t *d_words;
  if (!A.get(d_primary, 1, err) || !A.get(bc, 4 * nb, err) || !A.get(occ, 4 * nb, err) || !A.get(d_cnt, 20, err) ||
      !A.get(d_words, words, err))
    return false;
  DCK(hipMemsetAsync(d_cnt, 0, 20 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_find_primary<I>, dim3(grid_for(N)), dim3(kBlock), 0, s, SA, N, d_primary);
  uint8_t *bw = keep;  // n bytes, free now
  hipLaunchKernelGGL(k_bwt_bytes<I>, dim3(grid_for(n)), dim3(kBlock), 0, s, T, SA, d_primary, n, bw);
  hipLaunchKernelGGL(k_block_counts, dim3(grid_for(nb)), dim3(kBlock), 0, s, bw, n, nb, bc);
  for (int x = 0; x < 4; ++x) {
    size_t b = tmp_b;
    DCK(rocprim::exclusive_scan(tmp, b, bc + x * nb, occ + x * nb, (uint64_t)0, nb, rocprim::plus<uint64_t>(), s));
  }
  hipLaunchKernelGGL(k_bwt_words, dim3(grid_for(nb)), dim3(kBlock), 0, s, bw, occ, n, nb, d_words);
  hipLaunchKernelGGL(k_text_counts, dim3(grid_for(n)), dim3(kBlock), 0, s, T, n, d_cnt);
  DCK(hipGetLastError());
  const uint64_t n_sa = (n + 32) / 32;  // BWA sample: rows 32 i, i in [1, n_sa)
  uint64_t *d_sa32;
  if (!A.get(d_sa32, n_sa, err)) return false;
  if (n_sa > 1)
    hipLaunchKernelGGL((k_sample<I, uint64_t>), dim3(grid_for(n_sa - 1)), dim3(kBlock), 0, s, SA, (uint64_t)1,
                       (uint64_t)32, n_sa - 1, d_sa32);
  out.g3_intv = (uint32_t)dense_intv;
  out.g3_width = N < 0xffffffffULL ? 4 : 8;
  out.g3_count = (N + (uint64_t)dense_intv - 1) / (uint64_t)dense_intv;
  uint8_t *d_g3;
  if (!A.get(d_g3, out.g3_count * out.g3_width, err)) return false;
  if (out.g3_width == 4)
    hipLaunchKernelGGL((k_sample<I, uint32_t>), dim3(grid_for(out.g3_count)), dim3(kBlock), 0, s, SA, (uint64_t)0,
                       (uint64_t)dense_intv, out.g3_count, (uint32_t *)d_g3);
  else
    hipLaunchKernelGGL((k_sample<I, uint64_t>), dim3(grid_for(out.g3_count)), dim3(kBlock), 0, s, SA, (uint64_t)0,
                       (uint64_t)dense_intv, out.g3_count, (uint64_t *)d_g3);
  DCK(hipGetLastError());
  DCK(hipStreamSynchronize(s));
  progress("BWT and samples built on the device", 0, (int64_t)N);
  // ---- copy out
  out.n = n;
  DCK(hipMemcpy(&out.primary, d_primary, 8, hipMemcpyDeviceToHost));
  unsigned long long cnt[20];
  DCK(hipMemcpy(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost));
  out.bwt.resize((size_t)words);
  DCK(hipMemcpy(out.bwt.data(), d_words, (size_t)(nb * 16 > words ? words : nb * 16) * 4, hipMemcpyDeviceToHost));
  // the last block may hold fewer than 8 symbol words: its words were
  // written at their place, and the totals follow them
  uint64_t tot[4];
  for (int x = 0; x < 4; ++x) tot[x] = cnt[x];
  const uint64_t last_words = (n - 128 * (nb - 1) + 15) / 16;
  const uint64_t at = (nb - 1) * 16 + 8 + last_words;
  if (at + 8 != words) {
    err = "internal: .bwt word count";
    return false;
  }
  std::memcpy(out.bwt.data() + at, tot, sizeof tot);
  out.L2[0] = 0;
  for (int x = 0; x < 4; ++x) out.L2[x + 1] = out.L2[x] + tot[x];
  int64_t acc = 0;
  for (int x = 0; x < 5; ++x) {
    const int64_t v = x < 4 ? (int64_t)cnt[x] : 0;
    out.cnt1[x] = acc;
    acc += v;
  }
  acc = 0;
  for (int x = 0; x < 17; ++x) {
    const int64_t v = x < 16 ? (int64_t)cnt[4 + x] : 0;
    out.cnt2[x] = acc;
    acc += v;
  }
  out.sa.resize((size_t)(n_sa - 1));
  if (n_sa > 1) DCK(hipMemcpy(out.sa.data(), d_sa32, (size_t)(n_sa - 1) * 8, hipMemcpyDeviceToHost));
  out.g3sa.resize((size_t)(out.g3_count * out.g3_width));
  DCK(hipMemcpy(out.g3sa.data(), d_g3, out.g3sa.size(), hipMemcpyDeviceToHost));
  return true;
}

}  // namespace

bool build_fm_device(int device, const std::vector<uint8_t> &fwd, int dense_intv, FmFiles &out, std::string &err) {
  if (fwd.empty()) {
    err = "empty reference";
    return false;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    err = "no HIP device " + std::to_string(device);
    return false;
  }
  if (hipSetDevice(device) != hipSuccess) {
    err = "hipSetDevice failed";
    return false;
  }
  const uint64_t N = 2 * (uint64_t)fwd.size() + 1;
  if (N < 0xffffffffULL) return build_device<uint32_t>(fwd, dense_intv, out, err);
  return build_device<uint64_t>(fwd, dense_intv, out, err);
}

}  // namespace g3a
