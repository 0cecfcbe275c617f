// HIP kernels of the BWA-MEM hot path for MI355X (gfx950, wave64).
//
// Stage kernels (one launch each, all on the context's stream):
//   k_seed           thread per read: SMEM + reseed + third-round intervals
//   k_seed_slow      wavefront per read, for reads whose lists outgrew the
//                    fast path's buffers (LDS lists, list entries on lanes)
//   k_sa_lookup      thread per seed: SA interval row -> reference position
//   k_chain          wavefront per read: seed sort, chaining, chain filter,
//                    tasks (lane 0 in global memory for the rare large read)
//   k_task_compact   thread per read: per-read task lists -> dense task array
//   k_extend         one wavefront per task: banded local extension, left
//                    then right, anti-diagonal sweep over 64-row tiles
//   k_region_filter  thread per read
//   k_aln_prep       thread per read: kept regions -> alignment list
//   k_traceback      thread per alignment: global alignment + finalize
#include "g3a_dev.h"
#include "g3a_kernels.h"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace g3a {

// min(*dev, cap): a count known on the device, clamped to the capacity the
// launch was set up for (cap itself when dev is null)
__device__ __forceinline__ int64_t dcount(const int64_t *dev, int64_t cap) {
  if (!dev) return cap;
  const int64_t v = *dev;
  return v < cap ? v : cap;
}

// Register arrays of a wavefront: element e of an array of NC chunks lives
// in chunk e / 64, lane e % 64.
template <class T>
__device__ __forceinline__ T lane_read(T v, int l) {  // v of lane l (l wave-uniform)
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)v;
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return (T)((uint64_t)hi << 32 | lo);
  } else {
    return (T)__builtin_amdgcn_readlane((int)v, l);
  }
}
template <int NC, class T>
__device__ __forceinline__ T elem(const T (&v)[NC], int e) {  // e wave-uniform
  T x = lane_read(v[0], e & 63);
#pragma unroll
  for (int k = 1; k < NC; ++k)
    if ((e >> 6) == k) x = lane_read(v[k], e & 63);
  return x;
}
template <int NC, class T>
__device__ __forceinline__ void set_elem(T (&v)[NC], int e, T x, int lane) {  // e wave-uniform
#pragma unroll
  for (int k = 0; k < NC; ++k)
    if (e == 64 * k + lane) v[k] = x;
}
template <int NC, class T>
__device__ __forceinline__ T gather(const T (&v)[NC], int e) {  // e per lane; call it with every lane active
                                                                 // (a shuffle from an inactive lane gives 0)
  T x = __shfl(v[0], e & 63);
#pragma unroll
  for (int k = 1; k < NC; ++k) {
    const T y = __shfl(v[k], e & 63);
    if ((e >> 6) == k) x = y;
  }
  return x;
}
// The next item of a work counter for each active lane: one atomic per
// wavefront, its active lanes taking consecutive items.
__device__ __forceinline__ int wave_next(int32_t *ctr) {
  const uint64_t m = __ballot(true);
  const int leader = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)m) - 1);
  int base = 0;
  if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(ctr, __popcll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  return base + __popcll(m & ((1ull << (threadIdx.x & 63)) - 1));
}
__device__ __forceinline__ void wave_lds_sync() {  // this wavefront's LDS stores are seen by its lanes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// ---------------------------------------------------------------------------
// seeding
// ---------------------------------------------------------------------------
// Persistent grid: each thread takes the next read from a device counter
// (one atomic per wavefront and turn, its active lanes taking consecutive
// reads), so that a thread held up by a long read takes fewer
// (r = tid, tid + T, ... when no counter is given).  Its LEP
// lists live in a global pool slice of its own, contiguous per thread (a
// divergent list access then touches one cache line; private memory would
// scatter every entry over eight lines of the per-lane swizzle).
// In the narrow form the superblock counts (16 B per 2^21 rows, at most
// 32 KB) are copied to LDS, so an Occ lookup costs one 16-byte global load,
// and the first a.lep_lds entries of each thread's LEP list live in LDS
// after them (12 B per entry, lanes interleaved), the rest in the thread's
// global slice: the list traffic then stays out of the caches that the Occ
// lines need.
template <bool L, class B>
__device__ __forceinline__ void seed_reads(SeedArgs &a, B &buf, int gtid, int nthreads) {
  using I = typename B::I;
  int64_t steps = 0;
  // learned-index form: range lookups, lookups answered by the FM-index,
  // text runs, run bases, lookups the model narrowed, model windows missed
  int64_t lc[6] = {0, 0, 0, 0, 0, 0};
  for (int i = a.next_read ? wave_next(a.next_read) : gtid; i < a.n_reads;
       i = a.next_read ? wave_next(a.next_read) : i + nthreads) {
    const int r = a.order ? a.order[i] : i;
    const int off = a.off[r], len = a.off[r + 1] - off;
    const uint8_t *q = a.seq + off;
    IntvSink sink{a.intv + (int64_t)r * a.slot_cap, a.slot_cap, 0, 0, a.opt.max_occ};
    if (len > kMaxReadLen) {  // unaligned (documented limit)
      a.n_intv[r] = 0;
      a.n_seed[r] = 0;
      a.slow[r] = 0;
      continue;
    }
    bool ok = smem_round1<L>(a.idx, a.opt, q, len, sink, buf, &steps, lc);
    if (ok) {
      const int nsm = sink.n < sink.cap ? sink.n : sink.cap;
      if (sink.n <= sink.cap) {
        for (int k = 0; ok && k < nsm; ++k) {
          Intv smem = sink.slot[k];
          ok = reseed_one(a.idx, a.opt, q, len, smem, buf.a, buf.f, buf.cap, buf.cap_f, sink, &steps);
        }
      } else ok = false;
      if (ok) reseed_round3<decltype(I::s), L>(a.idx, a.opt, q, len, sink, &steps, lc);
    }
    ok = ok && sink.n <= sink.cap;
    a.slow[r] = ok ? 0 : 1;
    a.n_intv[r] = ok ? sink.n : 0;
    a.n_seed[r] = ok ? sink.n_seed : 0;
  }
  if (a.fm_steps) {  // one atomic per wavefront
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) steps += __shfl_xor(steps, o);
    if ((threadIdx.x & 63) == 0 && steps) atomicAdd((unsigned long long *)a.fm_steps, (unsigned long long)steps);
  }
  if (L && a.li_stats) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      int64_t v = lc[k];
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
      if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long *)a.li_stats + k, (unsigned long long)v);
    }
  }
}

// L: the learned-index form (text comparison of unique matches, the
// model's intervals in the last round); a separate kernel, so that the
// FM-only one keeps its register budget
template <class E, bool L>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(L ? 5 : 4))) k_seed(SeedArgs a) {
  extern __shared__ uint4 sb_lds[];
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nthreads = gridDim.x * blockDim.x;
  E *la = reinterpret_cast<E *>(a.lep_pool) + (size_t)gtid * (kLepCap + kFoundCap), *lf = la + kLepCap;
  if constexpr (E::kNarrow) {
    const int n_sb = (int)(((uint64_t)a.idx.n + 1) >> 21) + 1;
    const uint4 *src = (const uint4 *)a.idx.sbc;
    for (int k = threadIdx.x; k < n_sb; k += blockDim.x) sb_lds[k] = src[k];
    __syncthreads();
    a.idx.sbc = (const uint32_t *)sb_lds;
    LepBuf<SplitList, GlobalList<E>> buf{
        SplitList{(uint32_t *)(sb_lds + n_sb) + threadIdx.x, (int)blockDim.x, a.lep_lds, la}, GlobalList<E>{lf},
        kLepCap, kFoundCap};
    seed_reads<L>(a, buf, gtid, nthreads);
  } else {
    GlobalBuf<E> buf{GlobalList<E>{la}, GlobalList<E>{lf}, kLepCap, kFoundCap};
    seed_reads<false>(a, buf, gtid, nthreads);
  }
}

// Seeding order.  Overlapping reads of one locus walk the FM-index through
// the same intervals; given neighbouring lanes, their Occ lines are still in
// L2 when the second read needs them.  The key of a read is the smallest
// hash over its canonical 16-mers (both strands give the same key); sorting
// by it brings reads that share that 16-mer together.
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__global__ void __launch_bounds__(256) k_seed_keys(const uint8_t *seq, const int32_t *off, int n, uint32_t *key,
                                                   int32_t *idx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int o = off[r], len = off[r + 1] - o;
  ByteReader q(seq + o);
  uint32_t fw = 0, rv = 0, best = 0xffffffffu;
  int run = 0;
  for (int i = 0; i < len; ++i) {
    const uint32_t c = (uint32_t)q(i);
    if (c > 3) {
      run = 0;
      continue;
    }
    fw = fw << 2 | c;
    rv = rv >> 2 | (3u - c) << 30;
    if (++run >= 16) {
      const uint32_t h = mix32(fw < rv ? fw : rv);
      best = h < best ? h : best;
    }
  }
  key[r] = best >> 8;  // 24-bit key, sorted on bits [0, 24)
  idx[r] = r;
}

// The backward steps of rounds 1 and 2 on a wavefront: entry j on lane
// j % 64 of chunk j / 64 extends in parallel with the others, then one pass
// in list order over the lanes' results takes the decisions that depend on
// the entries before (which entries go on, the new start, the seed), and the
// entries that go on are written back in order.  The lists are in LDS and
// every lane sees the same list; the results equal SerialBack's.
struct WaveBack {
  static constexpr int kChunks = (kMaxReadLen + 63) / 64;  // no list of a read holds more entries
  template <class L, class LF>
  static __device__ int round1(const DevIndex &x, const Opt &o, const L &list, int np, int i, int c, const LF &fl,
                               int &nf, int cap, int cap_f, int &last_start, int64_t *) {
    using I = typename L::I;
    const int lane = threadIdx.x & 63;
    I ok[kChunks];
    int pm[kChunks], pn[kChunks];
    uint64_t ext[kChunks], fail[kChunks];
    if (np > 64 * kChunks) return -1;  // (no read of <= kMaxReadLen bases gets here)
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
      const int j = 64 * k + lane;
      const bool live = j < np;
      bool e = false;
      pm[k] = pn[k] = 0;
      if (live) {
        const I pj = list.get(j);
        pm[k] = pj.m;
        pn[k] = pj.n;
        if (c >= 0) {
          backward_ext_ks(x, pj, c, ok[k]);
          e = ok[k].s >= 1;
        }
      }
      ext[k] = __ballot(live && e);
      fail[k] = __ballot(live && !e);
    }
    int first_ext = np, first_fail = np;
#pragma unroll
    for (int k = kChunks - 1; k >= 0; --k) {
      if (ext[k]) first_ext = 64 * k + __ffsll((unsigned long long)ext[k]) - 1;
      if (fail[k]) first_fail = 64 * k + __ffsll((unsigned long long)fail[k]) - 1;
    }
    // a new start: the first entry that does not extend, none before it did
    if (first_fail < first_ext && i + 1 < last_start) {
      last_start = i + 1;
      I p = list.get(first_fail);
      if (p.n - i >= o.min_seed_len) {
        p.m = i + 1;
        if (nf == cap_f) return -1;
        fl.put(nf++, p);  // every lane stores the same entry
      }
    }
    // the entries that go on: each extension whose size differs from the
    // last one kept
    uint64_t keep[kChunks];
    int nc = 0;
    decltype(I::s) last_s = 0;
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
      keep[k] = 0;
      for (uint64_t m = ext[k]; m; m &= m - 1) {
        const int b = __ffsll((unsigned long long)m) - 1;
        const auto sb = lane_read(ok[k].s, b);
        if (nc == 0 || sb != last_s) {
          keep[k] |= 1ull << b;
          ++nc;
          last_s = sb;
        }
      }
    }
    if (nc > cap) return -1;
    wave_lds_sync();  // every entry was read above
    int before = 0;
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
      if ((keep[k] >> lane) & 1) {
        I e = ok[k];
        e.m = pm[k];
        e.n = pn[k];
        list.put(before + __popcll(keep[k] & ((1ull << lane) - 1)), e);
      }
      before += __popcll(keep[k]);
    }
    wave_lds_sync();
    return nc;
  }
  template <class L, class LF, class P>
  static __device__ int reseed(const DevIndex &x, const Opt &o, const L &leps, int num_old, int m, int base,
                               P min_seed_intv, const LF &seeds, int &ns, int cap_s, int64_t *) {
    using I = typename L::I;
    const int lane = threadIdx.x & 63;
    I ex[kChunks];
    int ln[kChunks];
    if (num_old > 64 * kChunks) return -1;  // (no read of <= kMaxReadLen bases gets here)
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
      const int j = 64 * k + lane;
      ln[k] = 0;
      if (j < num_old) {
        const I lep = leps.get(j);
        backward_ext_ks(x, lep, base, ex[k]);
        ex[k].m = m - 1;
        ex[k].n = lep.n;
        ln[k] = lep.n - lep.m + 1;
      }
    }
    P curr = min_seed_intv;
    bool collected = false;
    int nl = 0;
    uint64_t keep[kChunks];
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
      keep[k] = 0;
      const int n = num_old - 64 * k < 64 ? num_old - 64 * k : 64;
      for (int b = 0; b < n; ++b) {
        const P sb = lane_read(ex[k].s, b);
        if (sb >= curr) {
          keep[k] |= 1ull << b;
          ++nl;
          curr = sb;
        } else if (!collected && lane_read(ln[k], b) >= o.min_seed_len) {
          if (ns == cap_s) return -1;
          seeds.put(ns++, leps.get(64 * k + b));  // every lane stores the same entry
          collected = true;
        }
      }
    }
    wave_lds_sync();  // every entry was read above
    int before = 0;
#pragma unroll
    for (int k = 0; k < kChunks; ++k) {
      if ((keep[k] >> lane) & 1) leps.put(before + __popcll(keep[k] & ((1ull << lane) - 1)), ex[k]);
      before += __popcll(keep[k]);
    }
    wave_lds_sync();
    return nl;
  }
};

// Slow path for the reads listed in `list`, one wavefront per read: LEP
// lists of kSlowLep entries in LDS, which no read of <= 320 bases can
// outgrow, the backward steps on the wavefront's lanes (WaveBack) and the
// rest of the search on every lane alike.  Pass 0 counts and keeps the first
// 2 * kSlowLep intervals in the read's scratch; pass 1 puts all intervals at
// the pool offsets computed from the counts, copying them from the scratch
// when it still holds them all (`from_scratch`: every slow read had its own
// scratch in pass 0) and searching again otherwise.
__global__ void __launch_bounds__(64) k_seed_slow(SeedArgs a, const int32_t *list, int n_list, const int32_t *n_dev,
                                                  int c0, Intv *scratch, const int64_t *pool_off, Intv *pool, int pass,
                                                  int from_scratch) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ LepWide lds_a[kSlowLep], lds_f[kSlowLep];
  const int t = blockIdx.x, lane = threadIdx.x;
  int nv = n_list;
  if (n_dev) {
    const int m = *n_dev - c0;
    nv = m < 0 ? 0 : m < n_list ? m : n_list;
  }
  if (t >= nv) return;  // the whole wavefront
  const int r = list[t];
  const int off = a.off[r], len = a.off[r + 1] - off;
  const uint8_t *q = a.seq + off;
  Intv *sc = scratch + (int64_t)t * kSlowScratch;
  if (pass && from_scratch && a.n_intv[r] <= 2 * kSlowLep) {
    const Intv *src = sc + 3 * kSlowLep;
    Intv *dst = pool + pool_off[t];
    for (int k = lane, n = a.n_intv[r]; k < n; k += 64) dst[k] = src[k];
    return;
  }
  Intv *slot = pass ? pool + pool_off[t] : sc + 3 * kSlowLep;
  const int cap = pass ? (int)(pool_off[t + 1] - pool_off[t]) : 2 * kSlowLep;
  IntvSink sink{slot, cap, 0, 0, a.opt.max_occ};  // every lane pushes the same intervals
  GlobalBuf<LepWide> buf{GlobalList<LepWide>{lds_a}, GlobalList<LepWide>{lds_f}, kSlowLep, kSlowLep};
  if (!smem_round1<true, GlobalBuf<LepWide>, WaveBack>(a.idx, a.opt, q, len, sink, buf, nullptr)) {
    sink.n = 0;
    sink.n_seed = 0;
    smem_round1_per_end(a.idx, a.opt, q, len, sink, nullptr);
  }
  const int nsm = sink.n;
  for (int k = 0; k < nsm; ++k) {
    const Intv smem = slot[k];
    reseed_one<GlobalList<LepWide>, GlobalList<LepWide>, WaveBack>(a.idx, a.opt, q, len, smem, buf.a, buf.f,
                                                                    kSlowLep, kSlowLep, sink, nullptr);
  }
  reseed_round3<int64_t>(a.idx, a.opt, q, len, sink, nullptr);
  if (lane == 0) {
    a.n_intv[r] = sink.n;
    a.n_seed[r] = sink.n_seed;
  }
}

// ---------------------------------------------------------------------------
// SA lookup (saLookup, bwa/src/aux.cu:1365-1425): thread per seed
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_sa_lookup(SaArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= dcount(a.total_dev, a.total)) return;
  // read owning seed s: last r with seed_off[r] <= s
  int lo = 0, hi = a.n_reads;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a.seed_off[mid] <= s) lo = mid;
    else hi = mid;
  }
  const int r = lo;
  const int64_t local = s - a.seed_off[r];
  const Intv *iv = a.intv + a.intv_base[r];
  const int ni = a.n_intv[r];
  // interval owning the local seed index (intervals carry cumulative counts)
  int64_t acc = 0;
  int k = 0;
  const int64_t max_occ = a.opt.max_occ;
  for (; k < ni; ++k) {
    int64_t c = iv[k].s > max_occ ? max_occ : iv[k].s;
    if (local < acc + c) break;
    acc += c;
  }
  const Intv p = iv[k];
  const int64_t j = local - acc;
  const int64_t step = p.s > max_occ ? p.s / max_occ : 1;
  Seed out;
  const int64_t row = p.k + step * j;
  out.rbeg = a.idx.saf    ? (int64_t)a.idx.saf[row]
             : a.idx.occn ? sa_value_n(a.idx, (uint32_t)row)
                          : sa_value(a.idx, row, nullptr);
  out.qbeg = p.m;
  out.len = p.n - p.m + 1;
  out.rid = intv2rid(a.idx, out.rbeg, out.rbeg + out.len);
  out.pad = 0;
  a.seeds[s] = out;
}

// ---------------------------------------------------------------------------
// Seed sort + chaining + chain filter + per-read task list
// (BTreeChaining and CHAINFILTERING, bwa/src/aux.cu:1755-1988; the seed
// order of sortSeeds, aux.cu:1696-1752)
// ---------------------------------------------------------------------------

// One read on one thread from its sorted seeds `s`: the B-tree chaining, the
// filter and the read's tasks, with the temporaries at the pointers given
// (global slices).
__device__ __forceinline__ void chain_read(const ChainArgs &a, int r, int64_t b, const Seed *s, int ns, ChainTmp *ch,
                                           int32_t *nxt, int32_t *order, KbNode *nodes, int cap_nodes,
                                           uint16_t *w16, int32_t *tmpi, uint8_t *kept) {
  const int l_seq = a.off[r + 1] - a.off[r];
  int nch = chain_build(a.idx, a.opt, s, ns, ch, nxt, nodes, cap_nodes, order);
  if (nch < 0) {  // cannot happen with the node counts given; reported as an error
    atomicOr(a.error, 1);
    a.n_task[r] = 0;
    return;
  }
  int nk = chain_filter(a.opt, s, nxt, ch, order, nch, w16, tmpi, kept);
  int t = 0;
  Task *out = a.tasks_tmp + b;
  for (int c = 0; c < nk; ++c) {
    const ChainTmp &cc = ch[order[c]];
    int64_t r0, r1;
    int rid;
    chain_window(a.idx, a.opt, s, nxt, cc, l_seq, r0, r1, rid);
    for (int j = cc.head; j >= 0; j = nxt[j]) {
      Task tk;
      tk.rbeg = s[j].rbeg;
      tk.rmax0 = r0;
      tk.rmax1 = r1;
      tk.qbeg = s[j].qbeg;
      tk.len = s[j].len;
      tk.read = r;
      tk.rid = rid;
      out[t++] = tk;
    }
  }
  a.n_task[r] = t;
}

// One read on one thread, in global memory: sort_seeds, then chain_read.
// The wavefront kernel below hands a read here when it has more seeds than
// its registers hold, or when its B-tree would differ from a sorted array
// (see there).
__device__ void chain_one(const ChainArgs &a, int r) {
  const int64_t b = a.seed_off[r];
  const int ns = (int)(a.seed_off[r + 1] - b);
  a.n_task[r] = 0;
  if (ns == 0) return;
  Seed *s = a.seeds + b;
  if (ns > 1) sort_seeds(s, ns, a.sort_tmp + b);
  const int64_t nb0 = (b + 32 * (int64_t)r) / 4, nb1 = (a.seed_off[r + 1] + 32 * (int64_t)(r + 1)) / 4;
  chain_read(a, r, b, s, ns, a.ch + b, a.nxt + b, a.order + b, a.nodes + nb0, (int)(nb1 - nb0), a.w16 + b,
             a.tmpi + 2 * b, a.kept + b);
}

// Reads of at most kChainThread seeds (91% at 1 Gbp), one a thread, as
// chain_wave does them (the B-tree as a sorted array; a read that repeats a
// chain position past 11 chains runs chain_one), with everything in the
// thread's LDS slice: 32-bit words, word k of the lane at lds[64 k + lane]
// (no bank conflicts), byte arrays four entries a word.  The seeds: rbeg
// (two words), qbeg | len << 16, rid; a chain: head | tail << 8 | n << 16
// (its position and contig are its head seed's); the B-tree order: slot
// positions (two words) and chain ids; the filter: chain ids, weights and
// begin | end << 16 in weight order.
constexpr int kChainThread = 16;
// A thread's LDS slice, 104 words (6 blocks of 64 threads per compute unit).
// A seed's word at kRbHi holds bits 32..47 of its position (texts stay below
// 2^48 rows) over its contig id, or 0xffff for an id outside [0, 455): such
// seeds never start a chain (bwa/src/aux.cu:1799) and no chain's id equals
// the mark, so the comparisons of ids are unchanged.  The filter's arrays
// (kOc, kOw, kObe) take the place of the B-tree's positions (kTpLo, kTpHi),
// which the chaining no longer needs once they are written.
struct ChainLds {
  enum : int {
    kRbLo = 0, kRbHi = 16, kQl = 32,    // seeds: position bits 0..31; bits 32..47 | contig id; qbeg | len
    kNxt = 48,                          // bytes: next seed of the chain
    kCh = 52,                           // chains
    kTpLo = 68, kTpHi = 84, kTc = 100,  // B-tree order: positions, chain ids (bytes)
    kOc = 68, kOw = 72, kObe = 80,      // filter order: chain ids (bytes), weights (halves), begin | end
    kWords = 104
  };
  uint32_t *w;
  __device__ uint32_t get(int k) const { return w[64 * k]; }
  __device__ void set(int k, uint32_t v) const { w[64 * k] = v; }
  __device__ int getb(int k, int e) const { return (int)(int8_t)reinterpret_cast<const uint8_t *>(w + 64 * (k + (e >> 2)))[e & 3]; }
  __device__ void setb(int k, int e, int v) const { reinterpret_cast<uint8_t *>(w + 64 * (k + (e >> 2)))[e & 3] = (uint8_t)v; }
  __device__ int geth(int k, int e) const { return reinterpret_cast<const uint16_t *>(w + 64 * (k + (e >> 1)))[e & 1]; }
  __device__ void seth(int k, int e, int v) const { reinterpret_cast<uint16_t *>(w + 64 * (k + (e >> 1)))[e & 1] = (uint16_t)v; }
  __device__ int64_t rb(int i) const { return (int64_t)((uint64_t)(get(kRbHi + i) >> 16) << 32 | get(kRbLo + i)); }
  __device__ int rid(int i) const { return (int)(get(kRbHi + i) & 0xffff); }
  __device__ int qb(int i) const { return (int)(get(kQl + i) & 0xffff); }
  __device__ int ln(int i) const { return (int)(get(kQl + i) >> 16); }
  __device__ int64_t tp(int s) const { return (int64_t)((uint64_t)get(kTpHi + s) << 32 | get(kTpLo + s)); }
};
static_assert(kChainThread <= 16, "the slice's arrays hold 16 entries");
static_assert(ChainLds::kObe + 16 <= ChainLds::kTc && ChainLds::kTpHi + 16 <= ChainLds::kTc, "slice layout");

__device__ bool chain_thread(const ChainArgs &a, int r, int64_t b, int ns, const ChainLds &m) {
  const Opt &o = a.opt;
  const int64_t l_pac = a.idx.l_pac;
  // the seeds, sorted by insertion as they arrive (seed_less, stable)
  for (int i = 0; i < ns; ++i) {
    const Seed v = a.seeds[b + i];
    const uint32_t key = (uint32_t)v.qbeg | (uint32_t)v.len << 16;  // qbeg, len < 2^16
    int j = i;
    for (; j > 0; --j) {
      const uint32_t kj = m.get(ChainLds::kQl + j - 1);
      const bool less = (int)(key & 0xffff) < (int)(kj & 0xffff) ||
                        ((key & 0xffff) == (kj & 0xffff) && (int)(key >> 16) < (int)(kj >> 16));
      if (!less) break;
      m.set(ChainLds::kRbLo + j, m.get(ChainLds::kRbLo + j - 1));
      m.set(ChainLds::kRbHi + j, m.get(ChainLds::kRbHi + j - 1));
      m.set(ChainLds::kQl + j, kj);
    }
    const uint32_t rid16 = v.rid >= 0 && v.rid < 455 ? (uint32_t)v.rid : 0xffffu;
    m.set(ChainLds::kRbLo + j, (uint32_t)(uint64_t)v.rbeg);
    m.set(ChainLds::kRbHi + j, (uint32_t)((uint64_t)v.rbeg >> 32) << 16 | rid16);
    m.set(ChainLds::kQl + j, key);
  }
  for (int i = 0; i < ns; ++i) m.setb(ChainLds::kNxt, i, -1);
  // chaining (BTreeChaining, aux.cu:1755-1825)
  int nch = 0;
  bool dup = false;
  for (int i = 0; i < ns; ++i) {
    const int64_t prb = m.rb(i);
    const int pqb = m.qb(i), pln = m.ln(i), prid = m.rid(i);
    // the first slot whose position is >= prb
    int lo = 0, hi = nch;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (m.tp(mid) < prb) lo = mid + 1;
      else hi = mid;
    }
    const int n_lt = lo, first_eq = lo < nch && m.tp(lo) == prb ? lo : -1;
    if (nch) {
      const int low = first_eq >= 0 ? first_eq : n_lt - 1;
      if (low >= 0) {
        // test_and_merge (aux.cu:54-84) against chain c
        const int c = m.getb(ChainLds::kTc, low);
        const uint32_t cw = m.get(ChainLds::kCh + c);
        const int chd = (int)(cw & 0xff), ct = (int)((cw >> 8) & 0xff);
        const int64_t last_rb = m.rb(ct), first_rb = m.rb(chd);
        const int last_qb = m.qb(ct), last_ln = m.ln(ct), first_qb = m.qb(chd);
        const int64_t qend = last_qb + last_ln, rend = last_rb + last_ln;
        bool merged = false;
        if (prid == m.rid(chd)) {
          if (pqb >= first_qb && pqb + pln <= qend && prb >= first_rb && prb + pln <= rend) {
            merged = true;  // contained in the chain: dropped
          } else if (!((last_rb < l_pac || first_rb < l_pac) && prb >= l_pac)) {
            const int64_t xd = pqb - last_qb, yd = prb - last_rb;
            if (yd >= 0 && xd - yd <= o.w && yd - xd <= o.w && xd - last_ln < o.max_chain_gap &&
                yd - last_ln < o.max_chain_gap) {
              m.setb(ChainLds::kNxt, ct, i);
              m.set(ChainLds::kCh + c, (uint32_t)chd | (uint32_t)i << 8 | ((cw >> 16) + 1) << 16);
              merged = true;
            }
          }
        }
        if (merged) continue;
      }
    }
    if (prid < 0 || prid >= 455) continue;  // bwa/src/aux.cu:1799
    m.set(ChainLds::kCh + nch, (uint32_t)i | (uint32_t)i << 8 | 1u << 16);
    const int ins = first_eq >= 0 ? first_eq + 1 : n_lt;
    dup |= first_eq >= 0;
    for (int sl = nch; sl > ins; --sl) {
      m.set(ChainLds::kTpLo + sl, m.get(ChainLds::kTpLo + sl - 1));
      m.set(ChainLds::kTpHi + sl, m.get(ChainLds::kTpHi + sl - 1));
      m.setb(ChainLds::kTc, sl, m.getb(ChainLds::kTc, sl - 1));
    }
    m.set(ChainLds::kTpLo + ins, (uint32_t)(uint64_t)prb);
    m.set(ChainLds::kTpHi + ins, (uint32_t)((uint64_t)prb >> 32));
    m.setb(ChainLds::kTc, ins, nch);
    ++nch;
    if (dup && nch > kKbN) return false;  // a B-tree of several nodes with repeated keys
  }
  Task *out = a.tasks_tmp + b;
  if (nch == 0) {
    a.n_task[r] = 0;
    return true;
  }
  // chain filter (aux.cu:1828-1988): weights in B-tree order, then a stable
  // sort by decreasing weight
  for (int sl = 0; sl < nch; ++sl) {
    const int c = m.getb(ChainLds::kTc, sl);
    const uint32_t cw = m.get(ChainLds::kCh + c);
    const int chd = (int)(cw & 0xff), ct = (int)((cw >> 8) & 0xff);
    int64_t endq = 0, endr = 0;
    int wq = 0, wr = 0;
    for (int j = chd; j >= 0; j = m.getb(ChainLds::kNxt, j)) {
      const int q = m.qb(j), l = m.ln(j);
      const int64_t rbj = m.rb(j);
      if (q >= endq) wq += l;
      else if (q + l > endq) wq += (int)(q + l - endq);
      endq = endq > q + l ? endq : q + l;
      if (rbj >= endr) wr += l;
      else if (rbj + l > endr) wr += (int)(rbj + l - endr);
      endr = endr > rbj + l ? endr : rbj + l;
    }
    int wt = wr < wq ? wr : wq;
    wt = wt < 1 << 30 ? wt : (1 << 30) - 1;
    const int wv = (uint16_t)wt;
    const uint32_t be = (uint32_t)(uint16_t)m.qb(chd) | (uint32_t)(uint16_t)(m.qb(ct) + m.ln(ct)) << 16;
    int k = sl;
    for (; k > 0 && m.geth(ChainLds::kOw, k - 1) < wv; --k) {
      m.setb(ChainLds::kOc, k, m.getb(ChainLds::kOc, k - 1));
      m.seth(ChainLds::kOw, k, m.geth(ChainLds::kOw, k - 1));
      m.set(ChainLds::kObe + k, m.get(ChainLds::kObe + k - 1));
    }
    m.setb(ChainLds::kOc, k, c);
    m.seth(ChainLds::kOw, k, wv);
    m.set(ChainLds::kObe + k, be);
  }
  // the filter: chain i is dropped iff an earlier kept chain j overlaps it by
  // >= mask_level of the shorter and outweighs it (no ALT contigs)
  uint32_t kept = 0;
  for (int i = 0; i < nch; ++i) {
    const uint32_t bei = m.get(ChainLds::kObe + i);
    const uint16_t bi = (uint16_t)(bei & 0xffff), ei = (uint16_t)(bei >> 16);
    const uint16_t wi = (uint16_t)m.geth(ChainLds::kOw, i);
    bool drop = false;
    for (int j = 0; j < i && !drop; ++j) {
      if (!((kept >> j) & 1)) continue;
      const uint32_t bej = m.get(ChainLds::kObe + j);
      const uint16_t bj = (uint16_t)(bej & 0xffff), ej = (uint16_t)(bej >> 16);
      const int b_max = bj > bi ? bj : bi, e_min = ej < ei ? ej : ei;
      if (e_min > b_max) {
        const int li = ei - bi, lj = ej - bj, min_l = li < lj ? li : lj;
        if (e_min - b_max >= min_l * o.mask_level && min_l < o.max_chain_gap) {
          const uint16_t wj = (uint16_t)m.geth(ChainLds::kOw, j);
          drop = wi < wj * o.drop_ratio && wj - wi >= o.min_seed_len << 1;
        }
      }
    }
    if (!drop) kept |= 1u << i;
  }
  // tasks: kept chains in order, each chain's seeds in chain order with the
  // chain's reference window (ExtendingPairGenerate, aux.cu:2101-2192)
  const int l_seq = a.off[r + 1] - a.off[r];
  int t = 0;
  for (uint32_t km = kept; km; km &= km - 1) {
    const int c = m.getb(ChainLds::kOc, __ffs(km) - 1);
    const int chd = (int)(m.get(ChainLds::kCh + c) & 0xff);
    int64_t rmax0 = l_pac << 1, rmax1 = 0;
    for (int j = chd; j >= 0; j = m.getb(ChainLds::kNxt, j)) {
      const int q = m.qb(j), l = m.ln(j);
      const int64_t rbj = m.rb(j);
      const int64_t lo = rbj - (q + cal_max_gap(o, q));
      const int64_t hi = rbj + l + ((l_seq - q - l) + cal_max_gap(o, l_seq - q - l));
      rmax0 = rmax0 < lo ? rmax0 : lo;
      rmax1 = rmax1 > hi ? rmax1 : hi;
    }
    rmax0 = rmax0 > 0 ? rmax0 : 0;
    rmax1 = rmax1 < l_pac << 1 ? rmax1 : l_pac << 1;
    const int64_t mid = m.rb(chd);
    if (rmax0 < l_pac && l_pac < rmax1) {
      if (mid < l_pac) rmax1 = l_pac;
      else rmax0 = l_pac;
    }
    int is_rev;
    const int rid = pos2rid(a.idx, depos(a.idx, mid, is_rev));
    int64_t far_beg = a.idx.ctg_off[rid], far_end = far_beg + a.idx.ctg_len[rid];
    if (is_rev) {
      const int64_t tb = far_beg;
      far_beg = (l_pac << 1) - far_end;
      far_end = (l_pac << 1) - tb;
    }
    rmax0 = rmax0 > far_beg ? rmax0 : far_beg;
    rmax1 = rmax1 < far_end ? rmax1 : far_end;
    for (int j = chd; j >= 0; j = m.getb(ChainLds::kNxt, j)) {
      Task tk;
      tk.rbeg = m.rb(j);
      tk.rmax0 = rmax0;
      tk.rmax1 = rmax1;
      tk.qbeg = m.qb(j);
      tk.len = m.ln(j);
      tk.read = r;
      tk.rid = rid;
      out[t++] = tk;
    }
  }
  a.n_task[r] = t;
  return true;
}

__global__ void __launch_bounds__(64) k_chain_thread(ChainArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ uint32_t lds[64 * ChainLds::kWords];
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= a.n_reads) return;
  const int64_t b = a.seed_off[r];
  const int ns = (int)(a.seed_off[r + 1] - b);
  if (ns == 0 || ns > kChainThread) return;  // k_chain's
  const ChainLds m{lds + threadIdx.x};
  if (!chain_thread(a, r, b, ns, m)) chain_one(a, r);
}

constexpr int kChainWaveSeeds = 128;  // most seeds a read may have on the wavefront path
struct ChainStage {                   // one wavefront's LDS: the permutations of the sorts
  int64_t rbeg[kChainWaveSeeds];
  int32_t qbeg[kChainWaveSeeds], len[kChainWaveSeeds], rid[kChainWaveSeeds];
};

// One read of ns <= 64 * NC seeds on one wavefront.  Returns false, having
// written nothing, when the read must take chain_one instead.
//
// The reference's B-tree of chains (t = 6, up to 11 keys a node) is, while it
// has one node, a sorted array in which a key is inserted after the first
// equal key and the lookup returns the first equal key or else the last
// smaller one; with more nodes and distinct keys its lookups and in-order
// traversal are those of the sorted array.  Both hold here: chain positions
// sit in a lane array in that order, and a read whose chains repeat a
// position once there are more than 11 of them goes to chain_one.
template <int NC>
__device__ bool chain_wave(const ChainArgs &a, int r, int64_t b, int ns, int lane, ChainStage &st) {
  const Opt &o = a.opt;
  const int64_t l_pac = a.idx.l_pac;
  const Seed *sd = a.seeds + b;
  // ---- sort: seed_less, ties by input position (the stable order of sort_seeds)
  {
    Seed v[NC];
    uint32_t key[NC];
    int rank[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int e = 64 * k + lane;
      key[k] = 0xffffffffu;
      rank[k] = 0;
      if (e < ns) {
        v[k] = sd[e];
        key[k] = (uint32_t)v[k].qbeg << 16 | (uint32_t)v[k].len;  // qbeg, len < 2^16
      }
    }
    for (int i = 0; i < ns; ++i) {
      const uint32_t ki = elem<NC>(key, i);
#pragma unroll
      for (int k = 0; k < NC; ++k) rank[k] += ki < key[k] || (ki == key[k] && i < 64 * k + lane);
    }
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (64 * k + lane < ns) {
        st.rbeg[rank[k]] = v[k].rbeg;
        st.qbeg[rank[k]] = v[k].qbeg;
        st.len[rank[k]] = v[k].len;
        st.rid[rank[k]] = v[k].rid;
      }
  }
  wave_lds_sync();
  int64_t s_rb[NC];
  int s_qb[NC], s_ln[NC], s_rid[NC], s_nxt[NC], s_ch[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int e = 64 * k + lane;
    const bool in = e < ns;
    s_rb[k] = in ? st.rbeg[e] : 0;
    s_qb[k] = in ? st.qbeg[e] : 0;
    s_ln[k] = in ? st.len[e] : 0;
    s_rid[k] = in ? st.rid[e] : 0;
    s_nxt[k] = -1;
    s_ch[k] = -1;
  }
  wave_lds_sync();  // the staging is reused below
  // ---- chaining (BTreeChaining, aux.cu:1755-1825): chains by creation, and
  // their positions in B-tree order
  int64_t c_pos[NC], t_pos[NC];
  int c_rid[NC], c_n[NC], c_head[NC], c_tail[NC], t_cid[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    c_pos[k] = t_pos[k] = 0;
    c_rid[k] = c_n[k] = c_head[k] = c_tail[k] = t_cid[k] = 0;
  }
  int nch = 0;
  bool dup = false;
  for (int i = 0; i < ns; ++i) {
    const int64_t prb = elem<NC>(s_rb, i);
    const int pqb = elem<NC>(s_qb, i), pln = elem<NC>(s_ln, i), prid = elem<NC>(s_rid, i);
    int n_lt = 0, first_eq = -1;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const bool live = 64 * k + lane < nch;
      const uint64_t lt = __ballot(live && t_pos[k] < prb), eq = __ballot(live && t_pos[k] == prb);
      n_lt += __popcll(lt);
      if (first_eq < 0 && eq) first_eq = 64 * k + __ffsll((unsigned long long)eq) - 1;
    }
    if (nch) {
      const int low = first_eq >= 0 ? first_eq : n_lt - 1;
      if (low >= 0) {
        // test_and_merge (aux.cu:54-84) against chain c
        const int c = elem<NC>(t_cid, low);
        const int ct = elem<NC>(c_tail, c), chd = elem<NC>(c_head, c);
        const int64_t last_rb = elem<NC>(s_rb, ct), first_rb = elem<NC>(s_rb, chd);
        const int last_qb = elem<NC>(s_qb, ct), last_ln = elem<NC>(s_ln, ct), first_qb = elem<NC>(s_qb, chd);
        const int64_t qend = last_qb + last_ln, rend = last_rb + last_ln;
        bool merged = false;
        if (prid == elem<NC>(c_rid, c)) {
          if (pqb >= first_qb && pqb + pln <= qend && prb >= first_rb && prb + pln <= rend) {
            merged = true;  // contained in the chain: dropped
          } else if (!((last_rb < l_pac || first_rb < l_pac) && prb >= l_pac)) {
            const int64_t xd = pqb - last_qb, yd = prb - last_rb;
            if (yd >= 0 && xd - yd <= o.w && yd - xd <= o.w && xd - last_ln < o.max_chain_gap &&
                yd - last_ln < o.max_chain_gap) {
              set_elem<NC>(s_nxt, ct, i, lane);
              set_elem<NC>(s_ch, i, c, lane);
              set_elem<NC>(c_tail, c, i, lane);
              set_elem<NC>(c_n, c, elem<NC>(c_n, c) + 1, lane);
              merged = true;
            }
          }
        }
        if (merged) continue;
      }
    }
    if (prid < 0 || prid >= 455) continue;  // bwa/src/aux.cu:1799
    set_elem<NC>(c_pos, nch, prb, lane);
    set_elem<NC>(c_rid, nch, prid, lane);
    set_elem<NC>(c_n, nch, 1, lane);
    set_elem<NC>(c_head, nch, i, lane);
    set_elem<NC>(c_tail, nch, i, lane);
    set_elem<NC>(s_ch, i, nch, lane);
    // into the ordered positions: after the first equal one, else before
    // the first larger one; the slots above move up by one
    const int ins = first_eq >= 0 ? first_eq + 1 : n_lt;
    dup |= first_eq >= 0;
#pragma unroll
    for (int k = NC - 1; k >= 0; --k) {
      int64_t pp = __shfl_up(t_pos[k], 1);
      int pc = __shfl_up(t_cid[k], 1);
      if (k > 0 && lane == 0) {
        pp = lane_read(t_pos[k - 1], 63);
        pc = lane_read(t_cid[k - 1], 63);
      }
      const int e = 64 * k + lane;
      if (e > ins && e <= nch) {
        t_pos[k] = pp;
        t_cid[k] = pc;
      } else if (e == ins) {
        t_pos[k] = prb;
        t_cid[k] = nch;
      }
    }
    ++nch;
    if (dup && nch > kKbN) return false;  // a B-tree of several nodes with repeated keys
  }
  if (nch == 0) {
    if (lane == 0) a.n_task[r] = 0;
    return true;
  }
  // ---- chain filter (sortChainsDecreasingWeight + CHAINFILTERING_filter_kernel,
  // aux.cu:1828-1988): weights by a walk of each chain's seeds, one chain a lane
  int c_w[NC], c_b[NC], c_e[NC];
  {
    int j[NC], wq[NC], wr[NC];
    int64_t endq[NC], endr[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const bool live = 64 * k + lane < nch;
      j[k] = live ? c_head[k] : -1;
      wq[k] = wr[k] = 0;
      endq[k] = endr[k] = 0;
      c_b[k] = (uint16_t)gather<NC>(s_qb, live ? c_head[k] : 0);
      const int tl = live ? c_tail[k] : 0;
      c_e[k] = (uint16_t)(gather<NC>(s_qb, tl) + gather<NC>(s_ln, tl));
    }
    for (;;) {
      bool more = false;
#pragma unroll
      for (int k = 0; k < NC; ++k) more |= j[k] >= 0;
      if (!__any(more)) break;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const int jj = j[k] >= 0 ? j[k] : 0;
        const int q = gather<NC>(s_qb, jj), l = gather<NC>(s_ln, jj), nx = gather<NC>(s_nxt, jj);
        const int64_t rb = gather<NC>(s_rb, jj);
        if (j[k] >= 0) {
          if (q >= endq[k]) wq[k] += l;
          else if (q + l > endq[k]) wq[k] += (int)(q + l - endq[k]);
          endq[k] = endq[k] > q + l ? endq[k] : q + l;
          if (rb >= endr[k]) wr[k] += l;
          else if (rb + l > endr[k]) wr[k] += (int)(rb + l - endr[k]);
          endr[k] = endr[k] > rb + l ? endr[k] : rb + l;
          j[k] = nx;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      int w = wr[k] < wq[k] ? wr[k] : wq[k];
      w = w < 1 << 30 ? w : (1 << 30) - 1;
      c_w[k] = (uint16_t)w;  // the filter's 16-bit weights
    }
  }
  // B-tree order, then a stable sort by decreasing weight (all of a read's
  // chains: at most 128 here)
  int o_cid[NC], o_w[NC], o_be[NC];
  {
    int rank[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int cid = t_cid[k];
      o_cid[k] = cid;
      // every lane takes part in the shuffle: a lane that reads from one
      // outside a divergent branch gets 0
      const int wc = gather<NC>(c_w, cid);
      o_w[k] = 64 * k + lane < nch ? wc : -1;
      o_be[k] = (int)((uint32_t)gather<NC>(c_b, cid) | (uint32_t)gather<NC>(c_e, cid) << 16);
      rank[k] = 0;
    }
    for (int i = 0; i < nch; ++i) {
      const int wi = elem<NC>(o_w, i);
#pragma unroll
      for (int k = 0; k < NC; ++k) rank[k] += wi > o_w[k] || (wi == o_w[k] && i < 64 * k + lane);
    }
#pragma unroll
    for (int k = 0; k < NC; ++k)
      if (64 * k + lane < nch) {
        st.qbeg[rank[k]] = o_cid[k];
        st.len[rank[k]] = o_w[k];
        st.rid[rank[k]] = o_be[k];
      }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int e = 64 * k + lane;
      if (e < nch) {
        o_cid[k] = st.qbeg[e];
        o_w[k] = st.len[e];
        o_be[k] = st.rid[e];
      }
    }
    wave_lds_sync();
  }
  // the filter: chain i is dropped iff an earlier kept chain j overlaps it by
  // >= mask_level of the shorter and outweighs it (no ALT contigs)
  uint64_t kept[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) kept[k] = 0;
  for (int i = 0; i < nch; ++i) {
    const int be = elem<NC>(o_be, i), wi = elem<NC>(o_w, i);
    const uint16_t bi = (uint16_t)(be & 0xffff), ei = (uint16_t)((uint32_t)be >> 16);
    bool drop = false;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      bool hit = false;
      if (64 * k + lane < i && ((kept[k] >> lane) & 1)) {
        const uint16_t bj = (uint16_t)(o_be[k] & 0xffff), ej = (uint16_t)((uint32_t)o_be[k] >> 16);
        const int wj = o_w[k];
        const int b_max = bj > bi ? bj : bi, e_min = ej < ei ? ej : ei;
        if (e_min > b_max) {
          const int li = ei - bi, lj = ej - bj, min_l = li < lj ? li : lj;
          if (e_min - b_max >= min_l * o.mask_level && min_l < o.max_chain_gap)
            hit = (uint16_t)wi < (uint16_t)wj * o.drop_ratio && (uint16_t)wj - (uint16_t)wi >= o.min_seed_len << 1;
        }
      }
      drop |= __ballot(hit) != 0;
    }
    if (!drop) kept[i >> 6] |= 1ull << (i & 63);
  }
  // ---- tasks: kept chains in order, each chain's seeds in chain order
  // (ascending seed index) with the chain's reference window
  // (ExtendingPairGenerate, aux.cu:2101-2192)
  int64_t s_lo[NC], s_hi[NC];
  const int l_seq = a.off[r + 1] - a.off[r];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int q = s_qb[k], l = s_ln[k];
    s_lo[k] = s_rb[k] - (q + cal_max_gap(o, q));
    s_hi[k] = s_rb[k] + l + ((l_seq - q - l) + cal_max_gap(o, l_seq - q - l));
  }
  Task *out = a.tasks_tmp + b;
  int t = 0;
#pragma unroll
  for (int kk = 0; kk < NC; ++kk) {
    uint64_t m = kept[kk];
    while (m) {
      const int i = 64 * kk + __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      const int c = elem<NC>(o_cid, i);
      int64_t rmax0 = l_pac << 1, rmax1 = 0;
      for (int j = elem<NC>(c_head, c); j >= 0; j = elem<NC>(s_nxt, j)) {
        const int64_t lo = elem<NC>(s_lo, j), hi = elem<NC>(s_hi, j);
        rmax0 = rmax0 < lo ? rmax0 : lo;
        rmax1 = rmax1 > hi ? rmax1 : hi;
      }
      rmax0 = rmax0 > 0 ? rmax0 : 0;
      rmax1 = rmax1 < l_pac << 1 ? rmax1 : l_pac << 1;
      const int64_t mid = elem<NC>(s_rb, elem<NC>(c_head, c));
      if (rmax0 < l_pac && l_pac < rmax1) {
        if (mid < l_pac) rmax1 = l_pac;
        else rmax0 = l_pac;
      }
      int is_rev;
      const int rid = pos2rid(a.idx, depos(a.idx, mid, is_rev));
      int64_t far_beg = a.idx.ctg_off[rid], far_end = far_beg + a.idx.ctg_len[rid];
      if (is_rev) {
        const int64_t tb = far_beg;
        far_beg = (l_pac << 1) - far_end;
        far_end = (l_pac << 1) - tb;
      }
      rmax0 = rmax0 > far_beg ? rmax0 : far_beg;
      rmax1 = rmax1 < far_end ? rmax1 : far_end;
      int before = 0;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const uint64_t in = __ballot(s_ch[k] == c);
        if (s_ch[k] == c) {
          Task tk;
          tk.rbeg = s_rb[k];
          tk.rmax0 = rmax0;
          tk.rmax1 = rmax1;
          tk.qbeg = s_qb[k];
          tk.len = s_ln[k];
          tk.read = r;
          tk.rid = rid;
          out[t + before + __popcll(in & ((1ull << lane) - 1))] = tk;
        }
        before += __popcll(in);
      }
      t += before;
    }
  }
  if (lane == 0) a.n_task[r] = t;
  return true;
}

// Reads of more seeds than the wavefront path holds, listed so that k_chain
// starts them first: each runs on one lane for far longer than a wavefront
// read, and started in read order the last of them would be the kernel's
// tail.
__global__ void k_chain_big(const int32_t *stop, int n_reads, const int64_t *seed_off, int32_t *list,
                            int32_t *n_list) {
  if (stop && *stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_reads && seed_off[r + 1] - seed_off[r] > kChainWaveSeeds) list[atomicAdd(n_list, 1)] = r;
}

// One wavefront per read (four a block) for reads of more than kChainThread
// and at most kChainWaveSeeds seeds.  Wavefront g first takes the g-th
// listed large read (chain_one on lane 0), then read g unless it is a listed
// or a small one; reads that chain_wave hands back also run chain_one on
// lane 0.
__global__ void __launch_bounds__(256) k_chain(ChainArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ ChainStage stage[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
  if (r >= a.n_reads) return;  // the whole wavefront
  if (a.serial) {  // every read by the B-tree code on one thread
    if (lane == 0) chain_one(a, r);
    return;
  }
  if (r < *a.n_big && lane == 0) chain_one(a, a.big[r]);
  const int64_t b = a.seed_off[r];
  const int ns = __builtin_amdgcn_readfirstlane((int)(a.seed_off[r + 1] - b));
  bool done = true;
  if (ns == 0) {
    if (lane == 0) a.n_task[r] = 0;
  } else if (ns <= kChainThread) {
    // k_chain_thread's
  } else if (ns <= 64) {
    done = chain_wave<1>(a, r, b, ns, lane, stage[wv]);
  } else if (ns <= kChainWaveSeeds) {
    done = chain_wave<2>(a, r, b, ns, lane, stage[wv]);
  }  // else: listed
  if (!done && lane == 0) chain_one(a, r);
}

__global__ void __launch_bounds__(128) k_task_compact(CompactArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const int64_t src = a.seed_off[r], dst = a.task_off[r];
  const int n = (int)(a.task_off[r + 1] - dst);
  for (int t = 0; t < n; ++t) a.tasks[dst + t] = a.tasks_tmp[src + t];
}

// ---------------------------------------------------------------------------
// banded local extension (ksw_extend_warp2, bwa/src/ksw.cu:814-1034)
//
// One wavefront per task.  Lane l owns row 64*T+l of tile T and sweeps the
// tile's anti-diagonals: at step d it fills column j = d - l.  The diagonal
// input comes from lane l-1 two steps back and the E input from lane l-1 one
// step back, both moved by DPP wave_shr:1; lane 0 takes them from an LDS row
// buffer that lane 63 filled during the previous tile.  The cell set and the
// position rule are those of the reference's 32-lane program (see g3a_dev.h
// and the oracle), which do not depend on the lane schedule.
// ---------------------------------------------------------------------------
constexpr int kQMax = kMaxReadLen + 2;

__device__ __forceinline__ int shr1(int src, int old) {
  return __builtin_amdgcn_update_dpp(old, src, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    int t = __shfl_xor(v, o);
    v = v > t ? v : t;
  }
  return v;
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    uint64_t t = __shfl_xor(v, o);
    v = v > t ? v : t;
  }
  return v;
}

struct ExtOut {
  int score, qle, tle;
};

// qbuf: query (qlen bytes) in LDS; tb_of(i): target base of row i.
template <class TB>
__device__ ExtOut extend_wave(const Opt &o, int qlen, const uint8_t *qbuf, int tlen, TB tb_of, int h0, int end_bonus,
                              int *Hbuf, int *Ebuf, int64_t *cells) {
  const int lane = threadIdx.x & 63;
  const int oe_del = o.o_del + o.e_del, oe_ins = o.o_ins + o.e_ins;
  int max_ = 0;
  for (int i = 0; i < 25; ++i) max_ = max_ > o.mat[i] ? max_ : o.mat[i];
  int w = o.w;
  {
    int max_ins = (int)((double)(qlen * max_ + end_bonus - o.o_ins) / o.e_ins + 1.);
    max_ins = max_ins > 1 ? max_ins : 1;
    w = w < max_ins ? w : max_ins;
    int max_del = (int)((double)(qlen * max_ + end_bonus - o.o_del) / o.e_del + 1.);
    max_del = max_del > 1 ? max_del : 1;
    w = w < max_del ? w : max_del;
  }
  const int Q = qlen;
  const int stride = kQMax + 2;
  // row buffer 0: the first row's inputs (H(-1,j-1), E = 0)
  for (int j = lane; j <= Q; j += 64) {
    int ht = h0 - oe_ins - (j - 1) * o.e_ins;
    Hbuf[j] = j == 0 ? h0 : (ht > 0 ? ht : 0);
    Ebuf[j] = 0;
  }
  __syncthreads();
  int best_h = h0, best_i = -1, best_j = -1;
  int64_t ncell = 0;
  const int ntile = (tlen + 63) >> 6;
  for (int T = 0; T < ntile; ++T) {
    const int cur = T & 1, nxt = cur ^ 1;
    const int *Hc = Hbuf + cur * stride, *Ec = Ebuf + cur * stride;
    int *Hn = Hbuf + nxt * stride, *En = Ebuf + nxt * stride;
    const int i = 64 * T + lane;
    const bool row_ok = i < tlen;
    const int tbase = row_ok ? tb_of(i) : 0;
    uint64_t srow = 0;  // packed scores of this row's target base against query bases 0..4
#pragma unroll
    for (int qc = 0; qc < 5; ++qc) srow |= (uint64_t)(uint8_t)o.mat[tbase * 5 + qc] << (8 * qc);
    const int lim = 32 * (i / 32 + 1) + w + 1;
    const int hb = h0 - (o.o_del + o.e_del * (i + 1));
    const int boundary = hb > 0 ? hb : 0;
    const bool quirk_row = (i & 31) == 0 && i >= 32;
    const int dlo = 64 * T - w > 0 ? 64 * T - w : 0;
    // values of the two steps before dlo: column -1 carries the row's left
    // boundary (lane 0 reaches it only at step -1), earlier columns are empty
    int H1 = dlo - 1 - lane == -1 ? boundary : 0, H2 = dlo - 2 - lane == -1 ? boundary : 0, E1 = 0, F = 0;
    const int dhi = Q + 62;
    for (int d = dlo; d <= dhi; ++d) {
      const int j = d - lane;
      const int di = d < Q ? d : Q;
      const int hd0 = shr1(H2, Hc[di]);
      const int ein0 = shr1(E1, Ec[di]);
      const bool in_q = j >= 0 && j < Q;
      const int qb = in_q ? qbuf[j] : 0;
      const bool valid = in_q && row_ok && j >= i - w && j <= lim && (i + j > 30 || j <= i + w + 1);
      const bool quirk = quirk_row && j > i + w + 1;
      int ht = h0 - oe_ins - (j - 1) * o.e_ins;
      const int hd = quirk ? (ht > 0 ? ht : 0) : hd0;
      const int ein = quirk ? 0 : ein0;
      const int sc = (int)(int8_t)(srow >> (8 * qb));
      const int M = hd ? hd + sc : 0;
      int h = M > ein ? M : ein;
      h = h > F ? h : F;
      int t1 = M - oe_del;
      t1 = t1 > 0 ? t1 : 0;
      int e = ein - o.e_del;
      e = e > t1 ? e : t1;
      int t2 = M - oe_ins;
      t2 = t2 > 0 ? t2 : 0;
      int f = F - o.e_ins;
      f = f > t2 ? f : t2;
      const int hcur = valid ? h : (j == -1 ? boundary : 0);
      const int ecur = valid ? e : 0;
      F = valid ? f : 0;
      if (valid) {
        ++ncell;
        if (h > best_h) { best_h = h; best_i = i; best_j = j; }
        else if (h == best_h && (i == best_i || best_i < 0)) { best_i = i; best_j = j; }
      }
      if (lane == 63) {
        if (in_q) { Hn[j + 1] = hcur; En[j] = ecur; }
        else if (j == -1) Hn[0] = hcur;
      }
      H2 = H1;
      H1 = hcur;
      E1 = ecur;
    }
    __syncthreads();
  }
  // combine lanes (ksw.cu:1008-1033 with the 32-lane residue rule)
  const int best = wave_max_i32(best_h);
  const uint64_t key = best_i >= 0 && best_h == best
                           ? ((uint64_t)(uint32_t)best << 42) | ((uint64_t)(0x1FFFFF - best_i) << 21) |
                                 (uint64_t)(uint32_t)best_j
                           : 0;
  const uint64_t kmax = wave_max_u64(key);
  const uint64_t covered = __ballot(best_i >= 0);
  int bi = -1, bj = -1;
  if (best > h0 || (((covered | (covered >> 32)) & 0xffffffffULL) == 0xffffffffULL)) {
    if (kmax) {
      bi = 0x1FFFFF - (int)((kmax >> 21) & 0x1FFFFF);
      bj = (int)(kmax & 0x1FFFFF);
    }
  }
  if (cells) *cells += ncell;
  ExtOut r;
  r.score = best;
  r.qle = bj + 1;
  r.tle = bi + 1;
  return r;
}

__global__ void __launch_bounds__(64) k_extend(ExtArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ uint8_t qbuf[kQMax];
  __shared__ int Hbuf[2 * (kQMax + 2)], Ebuf[2 * (kQMax + 2)];
  const int64_t t = blockIdx.x;
  if (t >= dcount(a.n_dev, a.n_tasks)) return;
  const Task tk = a.tasks[t];
  const int off = a.off[tk.read], l_seq = a.off[tk.read + 1] - off;
  const uint8_t *read = a.seq + off;
  const int lane = threadIdx.x;
  int64_t cells = 0;
  Region g;
  g.rid = tk.rid;
  g.seedlen0 = tk.len;
  g.secondary = 0;
  int score;
  // left extension: reversed read prefix against reversed reference
  {
    const int qlen = tk.qbeg, tlen = (int)(tk.rbeg - tk.rmax0);
    const int h0 = tk.len * a.opt.a;
    if (qlen > 0) {
      for (int j = lane; j < qlen; j += 64) qbuf[j] = read[qlen - 1 - j];
      __syncthreads();
      const int64_t rb = tk.rbeg;
      ExtOut e = extend_wave(
          a.opt, qlen, qbuf, tlen, [&](int i) { return text_base(a.idx, rb - 1 - i); }, h0, a.opt.pen_clip5, Hbuf,
          Ebuf, &cells);
      score = e.score;
      g.qb = tk.qbeg - e.qle;
      g.rb = tk.rbeg - e.tle;
      __syncthreads();
    } else {
      score = h0;
      g.qb = 0;
      g.rb = tk.rbeg;
    }
  }
  // right extension
  {
    const int qlen = l_seq - tk.qbeg - tk.len, tlen = (int)(tk.rmax1 - tk.rbeg - tk.len);
    const int h0 = score;
    if (qlen > 0) {
      for (int j = lane; j < qlen; j += 64) qbuf[j] = read[tk.qbeg + tk.len + j];
      __syncthreads();
      const int64_t re0 = tk.rbeg + tk.len;
      ExtOut e = extend_wave(
          a.opt, qlen, qbuf, tlen, [&](int i) { return text_base(a.idx, re0 + i); }, h0, a.opt.pen_clip3, Hbuf,
          Ebuf, &cells);
      score = e.score;
      g.qe = tk.qbeg + tk.len + e.qle;
      g.re = tk.rbeg + tk.len + e.tle;
    } else {
      g.qe = tk.qbeg + tk.len;
      g.re = tk.rbeg + tk.len;
    }
  }
  g.score = score;
  if (lane == 0) a.regs[t] = g;
  if (a.cells) {
    // per-lane cell counts summed across the wave, one atomic per task
    int64_t c = cells;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) atomicAdd((unsigned long long *)a.cells, (unsigned long long)c);
  }
}

// ---------------------------------------------------------------------------
// banded local extension, batched form: one side at a time, tasks binned by
// query length, lane engine (g3a_ext.h) for the tasks it admits and one
// wavefront per task (extend_wave) for the rest.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    int t = __shfl_xor(v, o);
    v = v < t ? v : t;
  }
  return v;
}

// Inputs of one side of task t (ExtendingPairGenerate, bwa/src/aux.cu:2101-2192;
// localExtending, aux.cu:2289-2386): left = reversed read prefix against the
// reversed reference before the seed, right = read suffix against the
// reference after the seed; right h0 = left score.
struct SideTask {
  int qlen, tlen, h0, end_bonus;
  const uint8_t *read;
  int qbeg, qend;
  int64_t rbeg, rend;
};

__device__ __forceinline__ SideTask side_task(const ExtSideArgs &a, int64_t t) {
  const Task tk = a.tasks[t];
  const int off = a.off[tk.read], l_seq = a.off[tk.read + 1] - off;
  SideTask st;
  st.read = a.seq + off;
  st.qbeg = tk.qbeg;
  st.qend = tk.qbeg + tk.len;
  st.rbeg = tk.rbeg;
  st.rend = tk.rbeg + tk.len;
  if (a.side == 0) {
    st.qlen = tk.qbeg;
    st.tlen = (int)(tk.rbeg - tk.rmax0);
    st.h0 = tk.len * a.opt.a;
    st.end_bonus = a.opt.pen_clip5;
  } else {
    st.qlen = l_seq - tk.qbeg - tk.len;
    st.tlen = (int)(tk.rmax1 - tk.rbeg - tk.len);
    st.h0 = a.left[t].score;
    st.end_bonus = a.opt.pen_clip3;
  }
  return st;
}

// Tiles of kTilePer * 256 elements for the classifying kernels: a tile
// counts its members in LDS and reserves its places with one global atomic a
// counter (a block of 256 doing so contended on the few counters).  The
// order of a list's entries is that of the reservations; its consumers
// compute each entry on its own.
constexpr int kTilePer = 16;

__global__ void __launch_bounds__(256) k_ext_classify(ExtSideArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ int32_t cnt[2], base[2];
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t tile = (int64_t)blockIdx.x * (256 * kTilePer);
  const int64_t n = dcount(a.n_dev, a.n_tasks);
  int n_lane = 0, slow_at[kTilePer];
#pragma unroll
  for (int i = 0; i < kTilePer; ++i) {
    const int64_t t = tile + i * 256 + threadIdx.x;
    bool lane = false, slow = false;
    uint32_t key = 0;  // 0 for tasks not on the lane engine and past the tasks
    if (t < n) {
      const SideTask st = side_task(a, t);
      if (st.qlen <= 0) {
        ExtRes r{st.h0, 0, 0, 0};
        a.res[t] = r;
      } else {
        const int w = ext_band_w(a.opt, st.qlen, st.end_bonus);
        lane = a.lane_opt_ok && ext_lane_ok(a.opt, st.qlen, w, st.h0);
        slow = !lane;
      }
      if (lane) key = (uint32_t)st.qlen;
    }
    if (t < a.n_tasks) a.key[t] = key;
    n_lane += lane;
    slow_at[i] = slow ? atomicAdd(&cnt[1], 1) : -1;
  }
  if (n_lane) atomicAdd(&cnt[0], n_lane);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (cnt[0]) atomicAdd(a.ctr, cnt[0]);
    base[1] = cnt[1] ? atomicAdd(a.ctr + 1, cnt[1]) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kTilePer; ++i)
    if (slow_at[i] >= 0) a.slow_list[base[1] + slow_at[i]] = (int32_t)(tile + i * 256 + threadIdx.x);
}

template <int R, bool SYM, bool PK>
__global__ void __launch_bounds__(64) k_ext_lane(ExtSideArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  constexpr int NW = R / 8;
  __shared__ uint32_t wt[6 * 64 * NW];
  const int lane = threadIdx.x;
  uint32_t *rb = a.rowbuf + (size_t)blockIdx.x * (kExtQCap * 64) + lane;
  const int n_lane = a.ctr[0];
  int64_t cells = 0;
  for (;;) {
    int chunk = 0;
    if (lane == 0) chunk = atomicAdd(a.ctr + 2, 1);
    chunk = __shfl(chunk, 0);
    if ((int64_t)chunk * 64 >= n_lane) break;
    const int64_t k = (int64_t)chunk * 64 + lane;
    const bool valid = k < n_lane;
    const int64_t t = valid ? a.sorted[k] : 0;
    SideTask st;
    int w = 0;
    if (valid) {
      st = side_task(a, t);
      w = ext_band_w(a.opt, st.qlen, st.end_bonus);
    }
    const int qmax = wave_max_i32(valid ? st.qlen : 0);
    const int wmin = wave_min_i32(valid ? w : 1 << 20);
    const int wmax = wave_max_i32(valid ? w : 0);
    if (valid) {
      // everything the accessors need, by value (a capture by reference
      // would move the kernel arguments to private memory)
      const uint8_t *rd = st.read;
      const uint8_t *pac = a.idx.pac;
      const int64_t l_pac = a.idx.l_pac;
      const int qbeg = st.qbeg, qend = st.qend;
      const int64_t rbeg = st.rbeg, rend = st.rend;
      ExtLaneOut r;
      if (a.side == 0)
        r = ext_lane<R, SYM, PK>(
            a.opt, st.qlen, [=](int j) { return (int)rd[qbeg - 1 - j]; }, st.tlen,
            [=](int i) { return text_base_p(pac, l_pac, rbeg - 1 - i); }, w, st.h0, qmax, wmin, wmax, rb, 64,
            wt + lane * NW, 64 * NW);
      else
        r = ext_lane<R, SYM, PK>(
            a.opt, st.qlen, [=](int j) { return (int)rd[qend + j]; }, st.tlen,
            [=](int i) { return text_base_p(pac, l_pac, rend + i); }, w, st.h0, qmax, wmin, wmax, rb, 64,
            wt + lane * NW, 64 * NW);
      ExtRes o{r.score, r.qle, r.tle, 0};
      a.res[t] = o;
      cells += r.cells;
    }
  }
  if (a.cells) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cells += __shfl_xor(cells, o);
    if (lane == 0 && cells) atomicAdd((unsigned long long *)a.cells, (unsigned long long)cells);
  }
}

__global__ void __launch_bounds__(64) k_ext_wave(ExtSideArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ uint8_t qbuf[kQMax];
  __shared__ int Hbuf[2 * (kQMax + 2)], Ebuf[2 * (kQMax + 2)];
  const int lane = threadIdx.x;
  const int n_slow = a.ctr[1];
  int64_t cells = 0;
  for (int e = blockIdx.x; e < n_slow; e += gridDim.x) {
    const int64_t t = a.slow_list[e];
    const SideTask st = side_task(a, t);
    const uint8_t *rd = st.read;
    if (a.side == 0)
      for (int j = lane; j < st.qlen; j += 64) qbuf[j] = rd[st.qbeg - 1 - j];
    else
      for (int j = lane; j < st.qlen; j += 64) qbuf[j] = rd[st.qend + j];
    __syncthreads();
    const DevIndex &x = a.idx;
    const int64_t rbeg = st.rbeg, rend = st.rend;
    ExtOut r;
    if (a.side == 0)
      r = extend_wave(
          a.opt, st.qlen, qbuf, st.tlen, [&](int i) { return text_base(x, rbeg - 1 - i); }, st.h0, st.end_bonus, Hbuf,
          Ebuf, &cells);
    else
      r = extend_wave(
          a.opt, st.qlen, qbuf, st.tlen, [&](int i) { return text_base(x, rend + i); }, st.h0, st.end_bonus, Hbuf,
          Ebuf, &cells);
    if (lane == 0) {
      ExtRes o{r.score, r.qle, r.tle, 0};
      a.res[t] = o;
    }
    __syncthreads();
  }
  if (a.cells) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cells += __shfl_xor(cells, o);
    if (lane == 0 && cells) atomicAdd((unsigned long long *)a.cells, (unsigned long long)cells);
  }
}

// Region of each task from its two sides (localExtending, aux.cu:2326-2370).
__global__ void __launch_bounds__(256) k_ext_finish(ExtFinishArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= dcount(a.n_dev, a.n_tasks)) return;
  const Task tk = a.tasks[t];
  const ExtRes l = a.left[t], r = a.right[t];
  Region g;
  g.rid = tk.rid;
  g.seedlen0 = tk.len;
  g.secondary = 0;
  g.qb = tk.qbeg - l.qle;
  g.rb = tk.rbeg - l.tle;
  g.qe = tk.qbeg + tk.len + r.qle;
  g.re = tk.rbeg + tk.len + r.tle;
  g.score = r.score;
  a.regs[t] = g;
}

// k-mer table: thread per k-mer code
__global__ void __launch_bounds__(256) k_kmer_table(DevIndex x, uint32_t *table) {
  const uint32_t code = blockIdx.x * blockDim.x + threadIdx.x;
  if (code >= (1u << (2 * kKmer))) return;
  uint32_t e[3];
  kmer_interval(x, code, e);
  table[3 * (size_t)code] = e[0];
  table[3 * (size_t)code + 1] = e[1];
  table[3 * (size_t)code + 2] = e[2];
}
void launch_kmer_table(const DevIndex &x, uint32_t *table, hipStream_t s) {
  hipLaunchKernelGGL(k_kmer_table, dim3((1u << (2 * kKmer)) / 256), dim3(256), 0, s, x, table);
}

// Occ lines built on the device from the BWA .bwt body: thread per line
// (occ_line_from_bwa).
__global__ void __launch_bounds__(256) k_occ_lines(const uint32_t *W, uint64_t n, uint64_t primary, uint32_t *wide,
                                                   uint32_t *narrow, uint32_t *sbc) {
  const uint64_t rows = n + 1;
  const uint64_t n_lines = (rows >> 6) + 1 + (narrow ? (rows >> 5) + 1 : 0);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t L = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; L < n_lines; L += stride)
    occ_line_from_bwa(W, n, primary, L, wide, narrow, sbc);
}
void launch_occ_lines(const uint32_t *W, uint64_t n, uint64_t primary, uint32_t *wide, uint32_t *narrow,
                      uint32_t *sbc, hipStream_t s) {
  const uint64_t lines = ((n + 1) >> 6) + 1 + (narrow ? ((n + 1) >> 5) + 1 : 0);
  const uint64_t blocks = (lines + 255) / 256;
  hipLaunchKernelGGL(k_occ_lines, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, s, W, n, primary,
                     wide, narrow, sbc);
}

// Input codes above 4 (anything but A, C, G, T) become 4.
__global__ void k_clamp_codes(uint8_t *seq, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && seq[i] > 4) seq[i] = 4;
}
void launch_clamp_codes(uint8_t *seq, int64_t n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_clamp_codes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, seq, n);
}

__global__ void k_fill_i32(int32_t *p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// region filter, alignment list, traceback + finalize
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) k_region_filter(RegionArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const int64_t b = a.task_off[r];
  const int n = (int)(a.task_off[r + 1] - b);
  a.n_aln[r] = n ? region_filter(a.opt, a.regs + b, n, a.flag + b, a.sec + b) : 0;
}

__global__ void __launch_bounds__(128) k_aln_prep(AlnPrepArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.n_reads) return;
  const int64_t src = a.task_off[r], dst = a.aln_off[r];
  const int n = (int)(a.aln_off[r + 1] - dst);
  for (int k = 0; k < n; ++k) {
    const Region g = a.regs[src + k];
    a.aln_reg[dst + k] = g;
    a.aln_read[dst + k] = r;
    a.aln_rank[dst + k] = k;
    const int lq = g.qe - g.qb;
    const int64_t rlen = g.re - g.rb;
    const int w = global_bw(a.opt, lq, rlen, g.score);
    a.z_size[dst + k] = aln_z_size(lq, rlen, w);
    a.cig_size[dst + k] = aln_cigar_cap(lq, rlen, w);
  }
}

// Alignment order for k_traceback: those that need the banded global
// alignment (w > 0) first, the ungapped ones (w = 0) last, so that the DP
// loops of a wavefront run with all lanes busy.
// Traceback classes by band: 0: 2w+2 <= 16, 1: <= 32, 2: wider (w > 0),
// 3: ungapped (w = 0).  The row buffers of class c < 2 fit an LDS ring of
// 16 << c columns per lane, so the smaller classes run at higher occupancy.
__device__ __forceinline__ int tb_class(const TbArgs &a, int64_t t) {
  const Region g = a.aln_reg[t];
  const int w = global_bw(a.opt, g.qe - g.qb, g.re - g.rb, g.score);
  return w == 0 ? 3 : 2 * w + 2 <= 16 ? 0 : 2 * w + 2 <= 32 ? 1 : 2;
}
// tb_ctr[0..2]: alignments of classes 0..2 (tiles of kTilePer * 256)
__global__ void __launch_bounds__(256) k_tb_count(TbArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ int32_t cnt[3];
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t tile = (int64_t)blockIdx.x * (256 * kTilePer);
  const int64_t n = dcount(a.n_dev, a.n_alns);
  int c3[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < kTilePer; ++i) {
    const int64_t t = tile + i * 256 + threadIdx.x;
    const int c = t < n ? tb_class(a, t) : 3;
#pragma unroll
    for (int k = 0; k < 3; ++k) c3[k] += c == k;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (c3[k]) atomicAdd(&cnt[k], c3[k]);
  __syncthreads();
  if (threadIdx.x < 3 && cnt[threadIdx.x]) atomicAdd(a.tb_ctr + threadIdx.x, cnt[threadIdx.x]);
}
// The list holds classes 0, 1, 2 in turn from the front and class 3 from the
// back; tb_ctr[4..7] are the placement cursors (tiles of kTilePer * 256).
__global__ void __launch_bounds__(256) k_tb_order(TbArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ int32_t cnt[4], base[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t tile = (int64_t)blockIdx.x * (256 * kTilePer);
  const int64_t n_alns = dcount(a.n_dev, a.n_alns);
  int cls[kTilePer], at[kTilePer];
#pragma unroll
  for (int i = 0; i < kTilePer; ++i) {
    const int64_t t = tile + i * 256 + threadIdx.x;
    cls[i] = t < n_alns ? tb_class(a, t) : -1;
    at[i] = cls[i] >= 0 ? atomicAdd(&cnt[cls[i]], 1) : -1;
  }
  __syncthreads();
  if (threadIdx.x < 4) base[threadIdx.x] = cnt[threadIdx.x] ? atomicAdd(a.tb_ctr + 4 + threadIdx.x, cnt[threadIdx.x]) : 0;
  __syncthreads();
  const int64_t base1 = a.tb_ctr[0], base2 = base1 + a.tb_ctr[1];
#pragma unroll
  for (int i = 0; i < kTilePer; ++i) {
    const int c = cls[i];
    if (c < 0) continue;
    const int r = base[c] + at[i];
    const int64_t pos = c == 0 ? r : c == 1 ? base1 + r : c == 2 ? base2 + r : n_alns - 1 - r;
    a.tb_order[pos] = (int32_t)(tile + i * 256 + threadIdx.x);
  }
}

// Persistent grids over the ordered list.  k_traceback_dp<RB> takes one
// class: its row buffers are an RB-column ring per lane in LDS (lanes
// interleaved), or for the widest class with 2w+2 > RB the thread's slice of
// the global pool (2 * kQMax words, contiguous per thread); k_traceback takes
// the ungapped class.
template <int RB>
__global__ void __launch_bounds__(64) k_traceback_dp(TbArgs a, int cls) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  __shared__ int32_t ring_h[RB * 64], ring_e[RB * 64];
  __shared__ uint4 zstage[64 * RB / 16];  // each lane's row of direction bytes, RB bytes
  const int lane = threadIdx.x;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + lane;
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  int32_t *eh_h = a.eh_pool + gtid * 2 * kQMax, *eh_e = eh_h + kQMax;
  int64_t e0 = 0;
  for (int k = 0; k < cls; ++k) e0 += a.tb_ctr[k];
  const int64_t e1 = e0 + a.tb_ctr[cls];
  for (int64_t e = e0 + gtid; e < e1; e += nthreads) {
    const int64_t t = a.tb_order[e];
    const int r = a.aln_read[t];
    const int off = a.off[r], l_seq = a.off[r + 1] - off;
    const Region g = a.aln_reg[t];
    const int w = global_bw(a.opt, g.qe - g.qb, g.re - g.rb, g.score);
    Aln out;
    if (RB == 64 && 2 * w + 1 <= 128)  // direction bytes and score from k_tb_wave_dp
      trace_and_finalize(a.idx, a.opt, a.seq + off, l_seq, g, a.aln_rank[t], EhArray{eh_h, eh_e},
                         a.z + a.z_off[t], a.cigar + a.cig_off[t], a.cig_off[t], out, a.tb_nw + t);
    else if (2 * w + 2 <= RB)
      trace_and_finalize(a.idx, a.opt, a.seq + off, l_seq, g, a.aln_rank[t],
                         EhRing<RB>{ring_h + lane, ring_e + lane, 64,
                                    reinterpret_cast<uint32_t *>(zstage + lane * (RB / 16))},
                         a.z + a.z_off[t], a.cigar + a.cig_off[t], a.cig_off[t], out);
    else
      trace_and_finalize(a.idx, a.opt, a.seq + off, l_seq, g, a.aln_rank[t], EhArray{eh_h, eh_e},
                         a.z + a.z_off[t], a.cigar + a.cig_off[t], a.cig_off[t], out);
    a.alns[t] = out;
  }
}

// Banded global DP (global_dp's recurrence, ksw_global2, DR-9) of the
// widest class's alignments with 64 (NH - 1) < 2w+1 <= 64 NH band columns:
// one wavefront per alignment, lane L on columns beg + L + 64 h (h < NH) of
// each row.  F depends only on the row's M values, so a row is one max-plus
// prefix scan over the columns (F_c = max(-inf - c e, max_{c' < c} M_c' - o
// - e (c - c'))), carried from one 64-column half to the next; E and the
// next row's H come from neighbouring columns, shifted by the band's start.
// Writes the direction bytes and the score (tb_nw) that k_traceback_dp<64>
// then walks; identical to global_dp on every cell.
template <int NH>
__device__ __forceinline__ int col_get(const int (&v)[NH], int c) {
  int r = __shfl(v[0], c & 63);
  if (NH > 1) {
    const int r1 = __shfl(v[NH - 1], c & 63);
    r = c >= 64 ? r1 : r;
  }
  return r;
}
template <int NH>
__global__ void __launch_bounds__(256) k_tb_wave_dp(TbArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int L = threadIdx.x & 63;
  const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t e0 = (int64_t)a.tb_ctr[0] + a.tb_ctr[1], e1 = e0 + a.tb_ctr[2];
  const Opt &o = a.opt;
  const int MINUS_INF = -0x40000000;
  const int oe_del = o.o_del + o.e_del, oe_ins = o.o_ins + o.e_ins;
  for (int64_t e = e0 + wid; e < e1; e += nwaves) {
    const int64_t t = a.tb_order[e];
    const Region g = a.aln_reg[t];
    const int qlen = g.qe - g.qb, tlen = (int)(g.re - g.rb);
    const int w = global_bw(o, qlen, tlen, g.score);
    if (2 * w + 1 > 64 * NH || 2 * w + 1 <= 64 * (NH - 1)) continue;  // another launch's alignment
    const int r = a.aln_read[t];
    const uint8_t *query = a.seq + a.off[r] + g.qb;
    uint8_t *z = a.z + a.z_off[t];
    const int zs = z_stride(qlen, w);
    TextReader tr(a.idx.pac, a.idx.l_pac);
    int hp[NH], ep[NH], beg_p = 0, n_p = 0;
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      hp[h] = 0;
      ep[h] = MINUS_INF;
    }
    for (int i = 0; i < tlen; ++i) {
      const int beg = i > w ? i - w : 0, end = i + w + 1 < qlen ? i + w + 1 : qlen, n = end - beg;
      const ScoreRow sr = score_row(o, tr(g.rb + i));
      int m[NH], pm[NH], ein[NH];
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int c = 64 * h + L, j = beg + c;
        const bool act = c < n;
        const int qj = act ? (int)query[j] : 4;
        int hd;
        if (i == 0) {
          hd = j == 0 ? 0 : -(o.o_ins + o.e_ins * j);
          ein[h] = MINUS_INF;
        } else {
          const int delta = beg - beg_p, kp = c - 1 + delta, ke = c + delta;
          const int hs = col_get<NH>(hp, kp < 0 ? 0 : kp), es = col_get<NH>(ep, ke > 64 * NH - 1 ? 64 * NH - 1 : ke);
          hd = kp < 0 ? (beg_p == 0 ? -(o.o_del + o.e_del * i) : MINUS_INF) : hs;
          ein[h] = ke >= n_p ? MINUS_INF : es;
        }
        m[h] = hd + sr(qj);
        pm[h] = act ? m[h] - oe_ins + o.e_ins * (c + 1) : MINUS_INF;
      }
      // inclusive max scan of each half, the second carried by the first's total
#pragma unroll
      for (int h = 0; h < NH; ++h) {
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int u = __shfl_up(pm[h], d);
          pm[h] = L >= d && u > pm[h] ? u : pm[h];
        }
      }
      int carry = MINUS_INF;
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        pm[h] = carry > pm[h] ? carry : pm[h];
        int ex = __shfl_up(pm[h], 1);
        ex = L == 0 ? carry : ex;
        ex = ex < MINUS_INF ? MINUS_INF : ex;
        carry = __shfl(pm[h], 63);
        const int c = 64 * h + L;
        const int f = ex - c * o.e_ins;
        int dir = m[h] >= ein[h] ? 0 : 1, hh = m[h] >= ein[h] ? m[h] : ein[h];
        dir = hh >= f ? dir : 2;
        hh = hh >= f ? hh : f;
        int tt = m[h] - oe_del, e2 = ein[h] - o.e_del;
        dir |= e2 > tt ? 4 : 0;
        e2 = e2 > tt ? e2 : tt;
        tt = m[h] - oe_ins;
        dir |= f - o.e_ins > tt ? 32 : 0;
        if (c < n) z[(int64_t)i * zs + c] = (uint8_t)dir;
        hp[h] = hh;
        ep[h] = e2;
      }
      beg_p = beg;
      n_p = n;
    }
    const int score = col_get<NH>(hp, qlen - 1 - beg_p);
    if (L == 0) a.tb_nw[t] = score;
  }
}

__global__ void __launch_bounds__(64) k_traceback(TbArgs a) {
  if (a.stop && *a.stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
  int32_t *eh_h = a.eh_pool + gtid * 2 * kQMax, *eh_e = eh_h + kQMax;
  const int64_t n_alns = dcount(a.n_dev, a.n_alns);
  for (int64_t e = a.tb_ctr[0] + a.tb_ctr[1] + a.tb_ctr[2] + gtid; e < n_alns; e += nthreads) {
    const int64_t t = a.tb_order[e];
    const int r = a.aln_read[t];
    const int off = a.off[r], l_seq = a.off[r + 1] - off;
    const Region g = a.aln_reg[t];
    Aln out;
    trace_and_finalize(a.idx, a.opt, a.seq + off, l_seq, g, a.aln_rank[t], EhArray{eh_h, eh_e}, a.z + a.z_off[t],
                       a.cigar + a.cig_off[t], a.cig_off[t], out);
    a.alns[t] = out;
  }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static inline unsigned blocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// 16-byte list entries when every SA row fits 32 bits
// the narrow layout is in use (texts below 2^32 rows, unless the context
// was set to the wide layout): decided by the index, not by the text size
static bool seed_narrow(const DevIndex &x) { return x.occn != nullptr && x.sbc != nullptr; }
static size_t seed_lds_bytes(const DevIndex &x) {
  return seed_narrow(x) ? (size_t)((((uint64_t)x.n + 1) >> 21) + 1) * 16 : 0;
}
int seed_grid_threads(int n_cu, const DevIndex &x, int block) {
  int per_cu = 0;
  hipError_t e = seed_narrow(x)
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_seed<LepNarrow, false>, block,
                                                                    seed_lds_bytes(x))
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_seed<LepWide, false>, block, 0);
  if (e != hipSuccess || per_cu < 1) per_cu = 1024 / block;
  return n_cu * per_cu * block;
}
void launch_seed(const SeedArgs &a, int n_threads, int blocks_per_cu, int block, hipStream_t s) {
  if (!a.n_reads) return;
  const int need = blocks(a.n_reads, block), have = n_threads / block;
  int nb = need < have ? need : have;
  if (nb < 1) nb = 1;
  if (seed_narrow(a.idx)) {
    void (*kern)(SeedArgs) = a.idx.saf ? k_seed<LepNarrow, true> : k_seed<LepNarrow, false>;
    // The compute unit's LDS (160 KB) is shared by the blocks it is to hold,
    // and each block keeps one copy of the superblock counts, so larger
    // blocks leave more room for list entries; the LEP entries per thread are
    // the most that still lets that many blocks be resident (a block that is
    // not resident would only start when another finishes, doubling the
    // kernel's time).
    const size_t sb = seed_lds_bytes(a.idx);
    int dev = 0, n_cu = 0, max_lds = 65536;
    if (hipGetDevice(&dev) == hipSuccess) {
      hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
      if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
          max_lds < 16384)
        max_lds = 65536;
    }
    int want = blocks_per_cu;
    if (want <= 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&want, kern, block, sb) != hipSuccess || want < 1))
      want = 1024 / block;
    size_t per_block = (size_t)160 * 1024 / (size_t)want - 1024;  // slack for allocation granules
    per_block = per_block < (size_t)max_lds ? per_block : (size_t)max_lds;
    int capl = per_block > sb ? (int)((per_block - sb) / ((size_t)block * 12)) : 0;
    capl = capl < kLepCap ? capl : kLepCap;
    int resident = 0;
    for (; capl > 0; --capl) {
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, block, sb + (size_t)capl * block * 12) !=
          hipSuccess)
        resident = 0;
      if (resident >= want) break;
    }
    // no more blocks than the compute units hold at once
    if (n_cu > 0 && resident > 0 && nb > resident * n_cu) nb = resident * n_cu;
    SeedArgs b = a;
    b.lep_lds = capl;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(block), sb + (size_t)capl * block * 12, s, b);
  } else {
    hipLaunchKernelGGL((k_seed<LepWide, false>), dim3(nb), dim3(block), 0, s, a);
  }
}
void launch_seed_slow(const SeedArgs &a, const int32_t *list, int n_list, const int32_t *n_dev, int c0, Intv *scratch,
                      const int64_t *pool_off, Intv *pool, int pass, int from_scratch, hipStream_t s) {
  if (n_list)
    hipLaunchKernelGGL(k_seed_slow, dim3(n_list), dim3(64), 0, s, a, list, n_list, n_dev, c0, scratch, pool_off, pool,
                       pass, from_scratch);
}
void launch_sa_lookup(const SaArgs &a, hipStream_t s) {
  if (a.total) hipLaunchKernelGGL(k_sa_lookup, dim3(blocks(a.total, 256)), dim3(256), 0, s, a);
}
void launch_chain(const ChainArgs &a, hipStream_t s) {
  if (!a.n_reads) return;
  if (a.serial) {
    hipLaunchKernelGGL(k_chain, dim3(blocks(a.n_reads, 4)), dim3(256), 0, s, a);
    return;
  }
  hipLaunchKernelGGL(k_chain_big, dim3(blocks(a.n_reads, 256)), dim3(256), 0, s, a.stop, a.n_reads, a.seed_off,
                     a.big, a.n_big);
  hipLaunchKernelGGL(k_chain_thread, dim3(blocks(a.n_reads, 64)), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_chain, dim3(blocks(a.n_reads, 4)), dim3(256), 0, s, a);
}
void launch_task_compact(const CompactArgs &a, hipStream_t s) {
  if (a.n_reads) hipLaunchKernelGGL(k_task_compact, dim3(blocks(a.n_reads, 128)), dim3(128), 0, s, a);
}
void launch_extend(const ExtArgs &a, hipStream_t s) {
  if (a.n_tasks) hipLaunchKernelGGL(k_extend, dim3((unsigned)a.n_tasks), dim3(64), 0, s, a);
}
void launch_ext_classify(const ExtSideArgs &a, hipStream_t s) {
  if (a.n_tasks) hipLaunchKernelGGL(k_ext_classify, dim3(blocks(a.n_tasks, 256 * kTilePer)), dim3(256), 0, s, a);
}
void launch_ext_lane(const ExtSideArgs &a, int strip_rows, int packed, int n_blocks, hipStream_t s) {
  if (!a.n_tasks) return;
  const bool sym = a.opt.o_del + a.opt.e_del == a.opt.o_ins + a.opt.e_ins;
  void (*k)(ExtSideArgs);
  if (strip_rows == 32)
    k = sym ? (packed ? k_ext_lane<32, true, true> : k_ext_lane<32, true, false>)
            : (packed ? k_ext_lane<32, false, true> : k_ext_lane<32, false, false>);
  else
    k = sym ? (packed ? k_ext_lane<16, true, true> : k_ext_lane<16, true, false>)
            : (packed ? k_ext_lane<16, false, true> : k_ext_lane<16, false, false>);
  hipLaunchKernelGGL(k, dim3(n_blocks), dim3(64), 0, s, a);
}
void launch_ext_wave(const ExtSideArgs &a, int n_blocks, hipStream_t s) {
  if (a.n_tasks) hipLaunchKernelGGL(k_ext_wave, dim3(n_blocks), dim3(64), 0, s, a);
}
void launch_ext_finish(const ExtFinishArgs &a, hipStream_t s) {
  if (a.n_tasks) hipLaunchKernelGGL(k_ext_finish, dim3(blocks(a.n_tasks, 256)), dim3(256), 0, s, a);
}
void launch_fill_i32(int32_t *p, int n, int32_t v, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_i32, dim3(blocks(n, 256)), dim3(256), 0, s, p, n, v);
}
void launch_region_filter(const RegionArgs &a, hipStream_t s) {
  if (a.n_reads) hipLaunchKernelGGL(k_region_filter, dim3(blocks(a.n_reads, 128)), dim3(128), 0, s, a);
}
void launch_aln_prep(const AlnPrepArgs &a, hipStream_t s) {
  if (a.n_reads) hipLaunchKernelGGL(k_aln_prep, dim3(blocks(a.n_reads, 128)), dim3(128), 0, s, a);
}
int traceback_grid_threads(int n_cu) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_traceback, 64, 0) != hipSuccess || per_cu < 1)
    per_cu = 16;
  return n_cu * per_cu * 64;
}
int64_t traceback_pool_words(int n_threads) { return (int64_t)n_threads * 2 * kQMax; }
void launch_traceback(const TbArgs &a, int n_threads, int n_cu, hipStream_t s) {
  if (!a.n_alns) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, a.tb_ctr, 8, 0);
  const dim3 gt((unsigned)((a.n_alns + 256 * kTilePer - 1) / (256 * kTilePer)));
  hipLaunchKernelGGL(k_tb_count, gt, dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_tb_order, gt, dim3(256), 0, s, a);
  // the pool has a slice for each of the n_threads threads
  auto dp_blocks = [&](const void *k) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0) != hipSuccess || per_cu < 1) per_cu = 4;
    int b = n_cu * per_cu;
    if (b > n_threads / 64) b = n_threads / 64;
    return b < 1 ? 1 : b;
  };
  hipLaunchKernelGGL(k_traceback_dp<16>, dim3((unsigned)dp_blocks((const void *)k_traceback_dp<16>)), dim3(64), 0, s,
                     a, 0);
  hipLaunchKernelGGL(k_traceback_dp<32>, dim3((unsigned)dp_blocks((const void *)k_traceback_dp<32>)), dim3(64), 0, s,
                     a, 1);
  // the widest class: band DP by wavefronts (2w+1 <= 64 and <= 128), then walks
  hipLaunchKernelGGL(k_tb_wave_dp<1>, dim3((unsigned)(n_cu * 8)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_tb_wave_dp<2>, dim3((unsigned)(n_cu * 8)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_traceback_dp<64>, dim3((unsigned)dp_blocks((const void *)k_traceback_dp<64>)), dim3(64), 0, s,
                     a, 2);
  const int64_t need = (a.n_alns + 63) / 64, have = n_threads / 64;
  hipLaunchKernelGGL(k_traceback, dim3((unsigned)(need < have ? need : have)), dim3(64), 0, s, a);
}


// ---------------------------------------------------------------------------
// small helpers: slow-read list, interval bases, scans
// ---------------------------------------------------------------------------
__global__ void k_collect_slow(int n, const int32_t *slow, int32_t *list, int32_t *count, int64_t *intv_base,
                               int slot_cap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  intv_base[r] = (int64_t)r * slot_cap;
  if (slow[r]) list[atomicAdd(count, 1)] = r;
}
__global__ void k_gather_slow_counts(int cap, const int32_t *n_dev, const int32_t *list, const int32_t *n_intv,
                                     int64_t *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = *n_dev < cap ? *n_dev : cap;
  if (t < cap) out[t] = t < n ? n_intv[list[t]] : 0;
}
__global__ void k_set_slow_base(int cap, const int32_t *n_dev, const int32_t *list, const int64_t *off, int64_t base,
                                int64_t *intv_base) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = *n_dev < cap ? *n_dev : cap;
  if (t < n) intv_base[list[t]] = base + off[t];
}
__global__ void k_zero_tail(int64_t *p, int64_t cap, const int64_t *n_dev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap && i >= dcount(n_dev, cap)) p[i] = 0;
}
// only the first count that does not fit is flagged: the kernels after it
// stop, so the counts checked later are not meaningful
__global__ void k_cap_check(const int64_t *total, const int32_t *total32, int64_t cap, int32_t *flag, int32_t bit) {
  if (*flag) return;
  const int64_t v = total ? *total : (int64_t)*total32;
  if (v > cap) *flag = bit;
}
__global__ void k_gather_totals(Gather16 g, int k, int64_t *dst) {
  const int i = threadIdx.x;
  if (i < k) dst[i] = *g.src[i];
}
__global__ void k_zero_first(int64_t *p) { *p = 0; }

// CIGAR compaction for the host copy: the traceback writes each alignment's
// two CIGARs into a slot sized for the worst case; the copy holds only the
// words used, alignment by alignment (reference-format CIGAR, then the
// global one), and the records point into it.
__global__ void __launch_bounds__(256) k_cig_words(const Aln *alns, int64_t cap, const int64_t *n_dev, int32_t *words) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap) words[i] = i < dcount(n_dev, cap) ? alns[i].n_cigar + alns[i].n_core : 0;
}
__global__ void __launch_bounds__(256) k_cig_gather(Aln *alns, int64_t cap, const int64_t *n_dev, const int32_t *stop,
                                                    const int64_t *coff, const uint32_t *cig, uint32_t *out) {
  if (stop && *stop) return;  // a capacity check before this kernel failed: the batch is redone
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= dcount(n_dev, cap)) return;
  Aln a = alns[i];
  const int64_t o = coff[i];
  for (int k = 0; k < a.n_cigar; ++k) out[o + k] = cig[a.cigar_off + k];
  for (int k = 0; k < a.n_core; ++k) out[o + a.n_cigar + k] = cig[a.core_off + k];
  a.cigar_off = o;
  a.core_off = o + a.n_cigar;
  alns[i] = a;
}

void launch_collect_slow(int n, const int32_t *slow, int32_t *list, int32_t *count, int64_t *intv_base, int slot_cap,
                         hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_collect_slow, dim3(blocks(n, 256)), dim3(256), 0, s, n, slow, list, count, intv_base,
                            slot_cap);
}
void launch_gather_slow_counts(int cap, const int32_t *n_dev, const int32_t *list, const int32_t *n_intv, int64_t *out,
                               hipStream_t s) {
  if (cap) hipLaunchKernelGGL(k_gather_slow_counts, dim3(blocks(cap, 256)), dim3(256), 0, s, cap, n_dev, list, n_intv, out);
}
void launch_set_slow_base(int cap, const int32_t *n_dev, const int32_t *list, const int64_t *off, int64_t base,
                          int64_t *intv_base, hipStream_t s) {
  if (cap)
    hipLaunchKernelGGL(k_set_slow_base, dim3(blocks(cap, 256)), dim3(256), 0, s, cap, n_dev, list, off, base, intv_base);
}
void launch_zero_tail(int64_t *p, int64_t cap, const int64_t *n_dev, hipStream_t s) {
  if (cap > 0) hipLaunchKernelGGL(k_zero_tail, dim3(blocks(cap, 256)), dim3(256), 0, s, p, cap, n_dev);
}
void launch_cap_check(const int64_t *total, int64_t cap, int32_t *flag, int32_t bit, hipStream_t s) {
  hipLaunchKernelGGL(k_cap_check, dim3(1), dim3(1), 0, s, total, (const int32_t *)nullptr, cap, flag, bit);
}
void launch_cap_check32(const int32_t *total, int64_t cap, int32_t *flag, int32_t bit, hipStream_t s) {
  hipLaunchKernelGGL(k_cap_check, dim3(1), dim3(1), 0, s, (const int64_t *)nullptr, total, cap, flag, bit);
}
void launch_gather_totals(const Gather16 &g, int k, int64_t *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_totals, dim3(1), dim3(64), 0, s, g, k, dst);
}

struct WidenI32 {
  __host__ __device__ int64_t operator()(int32_t v) const { return (int64_t)v; }
};

void launch_seed_keys(const uint8_t *seq, const int32_t *off, int n, uint32_t *key, int32_t *idx, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_seed_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, seq, off, n, key, idx);
}
// Keys are the top 24 bits of the hash: finer than needed to keep each
// 16-mer's reads together, and one radix pass cheaper than 32.  They are
// stored shifted down and sorted on bits [0, 24): with a begin bit above 0,
// rocPRIM 4.2's merge-sort path (the one it takes for mid-sized inputs) was
// measured to return duplicated values.
size_t seed_sort_tmp_bytes(int64_t n) {
  size_t b = 0;
  rocprim::radix_sort_pairs(nullptr, b, (const uint32_t *)nullptr, (uint32_t *)nullptr, (const int32_t *)nullptr,
                            (int32_t *)nullptr, (size_t)(n > 0 ? n : 1), 0, 24);
  return b + 256;
}
int seed_sort(const uint32_t *key_in, uint32_t *key_out, const int32_t *val_in, int32_t *val_out, int64_t n, void *tmp,
              size_t tmp_bytes, hipStream_t s) {
  if (n <= 0) return 0;
  return rocprim::radix_sort_pairs(tmp, tmp_bytes, key_in, key_out, val_in, val_out, (size_t)n, 0, 24, s) == hipSuccess
             ? 0
             : -1;
}

void launch_cig_words(const Aln *alns, int64_t cap, const int64_t *n_dev, int32_t *words, hipStream_t s) {
  if (cap > 0)
    hipLaunchKernelGGL(k_cig_words, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, alns, cap, n_dev, words);
}
void launch_cig_gather(Aln *alns, int64_t cap, const int64_t *n_dev, const int32_t *stop, const int64_t *coff,
                       const uint32_t *cig, uint32_t *out, hipStream_t s) {
  if (cap > 0)
    hipLaunchKernelGGL(k_cig_gather, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, alns, cap, n_dev, stop,
                       coff, cig, out);
}

size_t scan_tmp_bytes(int64_t n) {
  size_t a = 0, b = 0;
  int64_t *p = nullptr;
  int32_t *q = nullptr;
  rocprim::inclusive_scan(nullptr, a, p, p, (size_t)(n > 0 ? n : 1), rocprim::plus<int64_t>());
  rocprim::inclusive_scan(nullptr, b, rocprim::make_transform_iterator(q, WidenI32()), p, (size_t)(n > 0 ? n : 1),
                          rocprim::plus<int64_t>());
  return (a > b ? a : b) + 256;
}
void scan_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_first, dim3(1), dim3(1), 0, s, out);
  if (n > 0) rocprim::inclusive_scan(tmp, tmp_bytes, in, out + 1, (size_t)n, rocprim::plus<int64_t>(), s);
}
void scan_i32(const int32_t *in, int64_t *out, int64_t n, void *tmp, size_t tmp_bytes, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_first, dim3(1), dim3(1), 0, s, out);
  if (n > 0)
    rocprim::inclusive_scan(tmp, tmp_bytes, rocprim::make_transform_iterator(in, WidenI32()), out + 1, (size_t)n,
                            rocprim::plus<int64_t>(), s);
}

// Counting sort of the lane tasks by descending key (kernels of its own:
// no memset nodes when the batch is captured in a graph).  hist[0, 512):
// tasks per key; hist[512, 1024): each key's next free place in out.
constexpr int kExtKeys = 512;
__global__ void __launch_bounds__(256) k_ext_hist(const uint32_t *key, int64_t cap, const int64_t *n_dev,
                                                  int32_t *hist) {
  __shared__ int32_t h[kExtKeys];
  const int64_t n = dcount(n_dev, cap);
  for (int k = threadIdx.x; k < kExtKeys; k += blockDim.x) h[k] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t k = key[t];
    if (k > 0 && k < kExtKeys) atomicAdd(&h[k], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kExtKeys; k += blockDim.x)
    if (h[k]) atomicAdd(&hist[k], h[k]);
}
// one block of kExtKeys threads: start of key k = tasks with keys above k
__global__ void __launch_bounds__(kExtKeys) k_ext_bins(int32_t *hist) {
  __shared__ int32_t c[kExtKeys];
  const int k = threadIdx.x;
  c[k] = hist[kExtKeys - 1 - k];  // descending keys in ascending slots
  __syncthreads();
  for (int o = 1; o < kExtKeys; o <<= 1) {
    const int v = k >= o ? c[k - o] : 0;
    __syncthreads();
    c[k] += v;
    __syncthreads();
  }
  const int key = kExtKeys - 1 - k;
  hist[kExtKeys + key] = c[k] - hist[key];  // exclusive
}
// A block places a tile of 4096 tasks: ranks within the tile from LDS
// counters, then one global reservation per key present in the tile (few
// blocks, so the reservations hardly contend).
constexpr int kScatterPer = 16;  // tasks per thread
__global__ void __launch_bounds__(256) k_ext_scatter(const uint32_t *key, int64_t cap, const int64_t *n_dev,
                                                     int32_t *hist, int32_t *out) {
  __shared__ int32_t h[kExtKeys], base[kExtKeys];
  const int64_t tile = (int64_t)blockIdx.x * (256 * kScatterPer);
  const int64_t n = dcount(n_dev, cap);
  if (tile >= n) return;  // the whole block: past the tasks
  for (int k = threadIdx.x; k < kExtKeys; k += blockDim.x) h[k] = 0;
  __syncthreads();
  uint32_t k[kScatterPer];
  int r[kScatterPer];
#pragma unroll
  for (int i = 0; i < kScatterPer; ++i) {
    const int64_t t = tile + i * 256 + threadIdx.x;
    k[i] = t < n ? key[t] : 0;
    r[i] = k[i] > 0 && k[i] < kExtKeys ? atomicAdd(&h[k[i]], 1) : 0;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < kExtKeys; j += blockDim.x)
    if (h[j]) base[j] = atomicAdd(&hist[kExtKeys + j], h[j]);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScatterPer; ++i)
    if (k[i] > 0 && k[i] < kExtKeys) out[base[k[i]] + r[i]] = (int32_t)(tile + i * 256 + threadIdx.x);
}
void ext_sort(const uint32_t *key, int64_t cap, const int64_t *n_dev, int32_t *hist, int32_t *out, hipStream_t s) {
  launch_fill_i32(hist, 2 * kExtKeys, 0, s);
  if (cap <= 0) return;
  hipLaunchKernelGGL(k_ext_hist, dim3((unsigned)std::min<int64_t>(blocks(cap, 256), 2048)), dim3(256), 0, s, key, cap,
                     n_dev, hist);
  hipLaunchKernelGGL(k_ext_bins, dim3(1), dim3(kExtKeys), 0, s, hist);
  hipLaunchKernelGGL(k_ext_scatter, dim3(blocks(cap, 256 * kScatterPer)), dim3(256), 0, s, key, cap, n_dev, hist, out);
}

}  // namespace g3a
