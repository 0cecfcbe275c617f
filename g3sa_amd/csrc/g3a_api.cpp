// C-ABI implementation: context, HBM buffers and stage orchestration for the
// BWA-MEM hot path (the reference's device_alloc / memcpy_index /
// memcpy_input / bwa_align / memcpy_output, bwa/src/cuda_wrapper.cu and
// bwa/src/bwamem.cu:92-425).
//
// All per-batch intermediate arrays are sized from exact counts: each stage
// writes per-read counts, a device scan turns them into offsets and the
// total is read back to size the next stage's buffers (grow-only, kept
// across batches).
#include "../../include/g3a.h"
#include "g3a_index.h"
#include "g3a_kernels.h"
#include "g3a_layout.h"
#include "g3a_learned.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

using namespace g3a;

// The allocation-free entry point's workspace: one array per device that
// the library runs on, held until the process ends.  Host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device (scratch_address); the array's name is a host-side stand-in.
static __device__ __align__(256) uint8_t g3a_noalloc_scratch[G3A_NOALLOC_SCRATCH_BYTES];

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool view = false;  // a region of the static scratch array, not an allocation
};

}  // namespace

// Per-batch workspace of the pipeline.  A context owns one (grown as
// batches need, kept across batches); the allocation-free entry point runs
// on a second one whose buffers are views into the device's static scratch
// array.
struct Work {
  DevBuf lep_pool, eh_pool, tb_order, tb_ctr, tb_nw;
  DevBuf seed_key, seed_key2, seed_idx, seed_order, seed_sort_tmp;  // seeding order
  // seeding
  DevBuf intv, n_intv, n_seed, slow, intv_base, seed_off, counters, slow_list, slow_scratch, slow_n, slow_off;
  // seeds / chains
  DevBuf seeds, seeds2, ch, nxt, order, tmpi, w16, kept, nodes, tasks_tmp, n_task, task_off, chain_big;
  // tasks / regions
  DevBuf tasks, regs, flag, sec, n_aln, aln_off;
  // alignments
  DevBuf aln_reg, aln_read, aln_rank, z_size, cig_size, z_off, cig_off, z, cigar, alns;
  DevBuf cig_words, cig_coff, cig_out;  // compacted CIGARs for the host copy
  DevBuf scan_tmp, totals;
  // batched-form extension
  DevBuf ext_res[2], ext_key, ext_sorted, ext_slow, ext_ctr, ext_hist, ext_rowbuf;
};

// Capacities of one attempt at a batch: the buffers the pipeline writes are
// sized for these counts before the first kernel runs, and the kernels read
// the true counts from device memory (clamped to these), so the batch runs
// without host round trips.  A count that does not fit sets a bit of the
// overflow flag; the kernels after it stop, and the batch is run again with
// that capacity grown.
struct Caps {
  int64_t seeds = 0, tasks = 0, alns = 0, z = 0, cig = 0, used = 0, pool = 0;
  int slow = 0;
};

// Device index buffers shared by a context and its siblings: the store owns
// them and frees them when the last context holding it is destroyed (each
// context's own fields are then non-owning views).
struct IndexStore {
  int device = 0;
  std::vector<DevBuf> bufs;
  ~IndexStore() {
    hipSetDevice(device);
    for (DevBuf &b : bufs)
      if (b.p && !b.view) hipFree(b.p);
  }
};

// A bump allocator over the static scratch array, active during one
// allocation-free call.
struct ScratchArena {
  uint8_t *base = nullptr;
  size_t cap = 0, used = 0;
};

struct g3a_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  Opt opt{};
  HostIndex h;
  bool index_on_device = false;
  DevIndex idx{};
  int max_reads = 0;
  int64_t max_bases = 0;
  // index
  DevBuf occ, pac, sa, ctg_off, ctg_len, ctg_alt, kmer, occn, sbc;
  DevBuf ctg_names, ctg_name_off, log_tab;  // for the record formatter
  // batch text (names, qualities) and formatted records
  bool has_text = false;
  int text_n = -1;
  DevBuf names, name_off, quals, qual_off, txt_len, txt_off, txt;
  DevBuf txt_bound, stage_off, txt_stage, fmt_flag;  // the one-pass formatter's bounds, staging and overflow flag
  int64_t txt_bytes = -1;
  DevBuf mapq, md_len, md_off, md;  // per alignment, for g3a_memcpy_output
  // batch input
  int n_reads = 0;
  bool has_input = false, has_output = false;
  DevBuf seq, off;
  int seed_ordered = 0;  // seed reads in minimizer order (1) or input order (0)
  int seed_threads = 0, tb_threads = 0;  // resident maxima (the pools are sized for them)
  // k_seed grid: one 1024-thread block per compute unit keeps a single copy
  // of the superblock counts in LDS and the rest for list entries
  int seed_blocks_per_cu = 1;             // 0: the occupancy maximum
  int seed_block = 1024;                  // k_seed threads per block
  int chain_serial = 0;                   // chaining: 0 thread / wavefront tiers, 1 the B-tree code per read
  int force_wide = 0;  // the layout of texts of 2^32 rows or more, on any text
  // per-batch workspace: `own`, or `scratch_work` during and after an
  // allocation-free call (its views stay valid for g3a_memcpy_output)
  Work own, scratch_work;
  Work *w = &own;
  ScratchArena *arena = nullptr;  // set while an allocation-free call runs
  uint8_t *scratch = nullptr;     // this device's static scratch array
  DevBuf fmt_scan_tmp;            // scan storage of the formatter and output copy
  DevBuf pe_hist, pe_stat, pe_v, pe_out;  // paired-end statistics and pair choice
  // grids of the running call (sizes of the per-thread pools)
  int run_seed_threads = 0, run_tb_threads = 0, run_lane_blocks = 0, run_slow_chunk = 16384;
  // batched-form extension
  int form = G3A_FORM_BATCHED;
  int n_cu = 0;
  int strip_rows = 16, lane_blocks_per_cu = 8, wave_blocks_per_cu = 8;
  int ext_packed = 1;  // lane engine: two strip rows per register in plain columns
  // host staging: readbacks, and page-locked copies of each batch's input
  // and output so that the transfers run at DMA speed
  int64_t *pinned = nullptr;
  static constexpr size_t kStage = (size_t)16 << 20;
  uint8_t *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {};
  std::vector<uint8_t> out_buf;
  hipEvent_t ev[8] = {};
  // single-kernel timing: k_seed, lane engine (2 sides), one-wavefront
  // extension (2 sides or the plain form's k_extend), k_traceback
  hipEvent_t kev[12] = {};
  int kev_used[6] = {0, 0, 0, 0, 0, 0};
  float kms[4] = {0, 0, 0, 0};
  // last batch results
  // learned-index form: full SA, keys, model, build scratch
  DevBuf li_saf, li_keys, li_root, li_leaf, li_knot, li_leaf_a, li_short, li_ctr;
  bool li_built = false;
  uint32_t li_leaf_w = 0;
  g3a_learned_info li_info{};
  int64_t total_seeds = 0, total_tasks = 0, total_alns = 0, total_cig = 0, n_slow = 0;
  float ms[6] = {0, 0, 0, 0, 0, 0};
  // batch capacities: per read (seeds, tasks, alignments, CIGAR words used)
  // and per base (direction bytes, CIGAR words), raised by the batches seen
  double per_read[4] = {16, 6, 2, 32}, per_base[2] = {4, 4};
  int slow_cap = 1024;          // slow-path reads per batch before a rerun
  int64_t pool_per_slow = 256;  // their intervals, per slow read
  int64_t n_bases = 0;          // bases of the resident batch
  Caps caps{};                  // of the last batch
  uint64_t buf_gen = 0;         // bumped by every allocation (graphs hold addresses)
  int64_t overflow_reruns = 0, graph_captures = 0;
  // one graph per batch shape (batched or plain form, input seeding order)
  int use_graph = 0;
  bool graph_used = false;
  // the device index is shared (g3a_device_alloc_sibling): set on the
  // context that made the index and on each sibling
  std::shared_ptr<IndexStore> store;
  bool sibling = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::string gkey;
};

namespace {

int fail(g3a_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCK(call)                                                                              \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return fail(ctx, G3A_E_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

int ensure(g3a_ctx *ctx, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 64;
  if (b.bytes >= bytes) return 0;
  if (ctx->arena) {  // allocation-free call: carve a view out of the scratch
    ScratchArena &a = *ctx->arena;
    const size_t at = (a.used + 255) & ~(size_t)255;
    if (at > a.cap || bytes > a.cap - at)
      return fail(ctx, G3A_E_ARG,
                  "the batch needs more than the " + std::to_string(a.cap >> 20) +
                      " MiB scratch of the allocation-free entry point (use g3a_bwa_align)");
    b.p = a.base + at;
    b.bytes = bytes;
    b.view = true;
    a.used = at + bytes;
    return 0;
  }
  if (b.view) return fail(ctx, G3A_E_INTERNAL, "scratch view grown outside an allocation-free call");
  if (b.p) {
    HIPCK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 4;  // headroom so repeated batches do not reallocate
  HIPCK(hipMalloc(&b.p, want));
  b.bytes = want;
  ++ctx->buf_gen;
  return 0;
}

// Host <-> device copies of pageable memory through two page-locked
// staging slots: the CPU copy into (or out of) one slot overlaps the DMA of
// the other.  Both end with the stream synchronized.
int stage_h2d(g3a_ctx *ctx, void *dev, const void *src, size_t bytes) {
  const size_t K = g3a_ctx::kStage;
  for (size_t c = 0, i = 0; c < bytes; c += K, ++i) {
    const int k = (int)(i & 1);
    const size_t m = bytes - c < K ? bytes - c : K;
    if (i >= 2) HIPCK(hipEventSynchronize(ctx->stage_ev[k]));
    std::memcpy(ctx->stage[k], (const uint8_t *)src + c, m);
    HIPCK(hipMemcpyAsync((uint8_t *)dev + c, ctx->stage[k], m, hipMemcpyHostToDevice, ctx->stream));
    HIPCK(hipEventRecord(ctx->stage_ev[k], ctx->stream));
  }
  HIPCK(hipStreamSynchronize(ctx->stream));
  return 0;
}

int stage_d2h(g3a_ctx *ctx, void *dst, const void *dev, size_t bytes) {
  const size_t K = g3a_ctx::kStage, n = (bytes + K - 1) / K;
  auto issue = [&](size_t i) -> int {
    const size_t c = i * K, m = bytes - c < K ? bytes - c : K;
    HIPCK(hipMemcpyAsync(ctx->stage[i & 1], (const uint8_t *)dev + c, m, hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(hipEventRecord(ctx->stage_ev[i & 1], ctx->stream));
    return 0;
  };
  for (size_t i = 0; i < n && i < 2; ++i)
    if (int rc = issue(i)) return rc;
  for (size_t i = 0; i < n; ++i) {
    const size_t c = i * K, m = bytes - c < K ? bytes - c : K;
    HIPCK(hipEventSynchronize(ctx->stage_ev[i & 1]));
    std::memcpy((uint8_t *)dst + c, ctx->stage[i & 1], m);
    if (i + 2 < n)
      if (int rc = issue(i + 2)) return rc;
  }
  HIPCK(hipStreamSynchronize(ctx->stream));
  return 0;
}

template <class T>
T *P(DevBuf &b) {
  return (T *)b.p;
}

void free_buf(DevBuf &b) {
  if (b.p && !b.view) hipFree(b.p);
  b = DevBuf{};
}

// Device address of g3a_noalloc_scratch on `device` (current), looked up
// once per device and kept: each device holds its own copy of the array.
std::mutex g_scratch_m;
uint8_t *g_scratch_addr[64] = {};

hipError_t scratch_address(int device, uint8_t **out) {
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> l(g_scratch_m);
  if (!g_scratch_addr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(g3a_noalloc_scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = static_cast<uint8_t *>(ptr);
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}

// Stream-side set-up of a context (its stream already created): events,
// page-locked staging, the device's scratch array.
int init_runtime(g3a_ctx *ctx) {
  for (auto &ev : ctx->ev) hipEventCreate(&ev);
  for (auto &ev : ctx->kev) hipEventCreate(&ev);
  if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess ||
      ctx->n_cu < 1)
    ctx->n_cu = 256;
  if (hipHostMalloc((void **)&ctx->pinned, 64 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&ctx->stage[0], g3a_ctx::kStage, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&ctx->stage[1], g3a_ctx::kStage, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->stage_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->stage_ev[1], hipEventDisableTiming) != hipSuccess)
    return G3A_E_HIP;
  // the allocation-free entry point's scratch array on this device (set up
  // here, with the device current)
  if (scratch_address(ctx->device, &ctx->scratch) != hipSuccess) return G3A_E_HIP;
  return 0;
}

// The device index's buffers of a context (shared with its siblings).
std::vector<DevBuf *> index_bufs(g3a_ctx *c) {
  return {&c->occ,     &c->kmer,       &c->occn,    &c->sbc,     &c->pac,     &c->sa,       &c->ctg_off,
          &c->ctg_len, &c->ctg_alt,    &c->ctg_names, &c->ctg_name_off, &c->log_tab, &c->li_saf, &c->li_keys,
          &c->li_root, &c->li_leaf,    &c->li_knot, &c->li_leaf_a, &c->li_short, &c->li_ctr};
}

}  // namespace

static int reserve_workspace(g3a_ctx *ctx);

extern "C" {

void g3a_opt_init(g3a_opt *o) {
  std::memset(o, 0, sizeof *o);
  o->a = 1; o->b = 4;
  o->o_del = o->o_ins = 6;
  o->e_del = o->e_ins = 1;
  o->w = 100;
  o->T = 30;
  o->pen_clip5 = o->pen_clip3 = 5;
  o->max_mem_intv = 20;
  o->min_seed_len = 19;
  o->split_width = 10;
  o->max_occ = 500;
  o->max_chain_gap = 10000;
  o->mask_level = 0.50f;
  o->drop_ratio = 0.50f;
  o->split_factor = 1.5f;
  o->mask_level_redun = 0.95f;
  int k = 0;
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) o->mat[k++] = (int8_t)(i == j ? o->a : -o->b);
    o->mat[k++] = -1;
  }
  for (int j = 0; j < 5; ++j) o->mat[k++] = -1;
}

const char *g3a_last_error(const g3a_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int g3a_build_index(const char *fasta, const char *prefix, int sa_intv, char *err, int err_len) {
  std::string e;
  if (!fasta || !prefix) e = "null argument";
  else if (build_index(fasta, prefix, sa_intv, e)) return G3A_OK;
  if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.c_str());
  return e == "null argument" ? G3A_E_ARG : G3A_E_INDEX;
}

int g3a_build_index_dev(int device, const char *fasta, const char *prefix, int sa_intv, char *err, int err_len) {
  std::string e;
  int rc = G3A_E_INDEX;
  if (!fasta || !prefix || device < 0) {
    e = "invalid argument";
    rc = G3A_E_ARG;
  } else if (build_index(fasta, prefix, sa_intv, e, device)) {
    return G3A_OK;
  }
  if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", e.c_str());
  return rc;
}

int g3a_device_alloc(int device, const char *prefix, const g3a_opt *o, int max_reads, int64_t max_bases,
                     g3a_ctx **out) {
  if (!out || !prefix || max_reads < 1 || max_bases < 1) return G3A_E_ARG;
  *out = nullptr;
  g3a_ctx *ctx = new g3a_ctx;
  ctx->device = device;
  ctx->max_reads = max_reads;
  ctx->max_bases = max_bases;
  g3a_opt d;
  if (!o) { g3a_opt_init(&d); o = &d; }
  Opt &p = ctx->opt;
  p.a = o->a; p.b = o->b; p.o_del = o->o_del; p.e_del = o->e_del; p.o_ins = o->o_ins; p.e_ins = o->e_ins;
  p.w = o->w; p.T = o->T; p.pen_clip5 = o->pen_clip5; p.pen_clip3 = o->pen_clip3;
  p.min_seed_len = o->min_seed_len; p.split_width = o->split_width; p.max_occ = o->max_occ;
  p.max_chain_gap = o->max_chain_gap; p.flag = o->flag; p.max_mem_intv = o->max_mem_intv;
  p.split_factor = o->split_factor; p.mask_level = o->mask_level; p.drop_ratio = o->drop_ratio;
  p.mask_level_redun = o->mask_level_redun;
  std::memcpy(p.mat, o->mat, 25);
  if (p.max_occ < 1 || p.min_seed_len < 1 || p.e_del < 1 || p.e_ins < 1) {
    delete ctx;
    return G3A_E_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    delete ctx;
    return G3A_E_HIP;
  }
  std::string e;
  if (!load_index(prefix, ctx->h, e)) {
    delete ctx;
    return G3A_E_INDEX;
  }
  for (auto &c : ctx->h.anns)
    if (c.n_ambs < 0) { delete ctx; return G3A_E_INDEX; }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return G3A_E_HIP;
  }
  if (init_runtime(ctx)) {
    g3a_destroy(ctx);
    return G3A_E_HIP;
  }
  *out = ctx;
  return G3A_OK;
}

int g3a_device_alloc_sibling(g3a_ctx *src, int max_reads, int64_t max_bases, g3a_ctx **out) {
  if (!src || !out || max_reads < 1 || max_bases < 1) return G3A_E_ARG;
  *out = nullptr;
  if (!src->index_on_device) return fail(src, G3A_E_STATE, "the source context has no index on the device");
  if (hipSetDevice(src->device) != hipSuccess) return G3A_E_HIP;
  g3a_ctx *ctx = new g3a_ctx;
  ctx->device = src->device;
  ctx->max_reads = max_reads;
  ctx->max_bases = max_bases;
  ctx->opt = src->opt;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return G3A_E_HIP;
  }
  if (init_runtime(ctx)) {
    g3a_destroy(ctx);
    return G3A_E_HIP;
  }
  // the index buffers the source still owns move into the shared store
  if (!src->store) {
    src->store = std::make_shared<IndexStore>();
    src->store->device = src->device;
  }
  const std::vector<DevBuf *> from = index_bufs(src), to = index_bufs(ctx);
  for (DevBuf *b : from)
    if (b->p && !b->view) {
      src->store->bufs.push_back(*b);
      b->view = true;
    }
  for (size_t i = 0; i < from.size(); ++i) {
    *to[i] = *from[i];
    to[i]->view = true;
  }
  ctx->store = src->store;
  ctx->sibling = true;
  ctx->index_on_device = true;
  ctx->idx = src->idx;
  ctx->force_wide = src->force_wide;
  ctx->li_built = src->li_built;
  ctx->li_leaf_w = src->li_leaf_w;
  ctx->li_info = src->li_info;
  // the source's form and tuning
  ctx->form = src->form;
  ctx->strip_rows = src->strip_rows;
  ctx->lane_blocks_per_cu = src->lane_blocks_per_cu;
  ctx->wave_blocks_per_cu = src->wave_blocks_per_cu;
  ctx->ext_packed = src->ext_packed;
  ctx->seed_blocks_per_cu = src->seed_blocks_per_cu;
  ctx->seed_block = src->seed_block;
  ctx->chain_serial = src->chain_serial;
  ctx->seed_ordered = src->seed_ordered;
  ctx->seed_threads = src->seed_threads;
  ctx->tb_threads = src->tb_threads;
  ctx->use_graph = src->use_graph;
  std::memcpy(ctx->per_read, src->per_read, sizeof ctx->per_read);
  std::memcpy(ctx->per_base, src->per_base, sizeof ctx->per_base);
  ctx->slow_cap = src->slow_cap;
  ctx->pool_per_slow = src->pool_per_slow;
  if (reserve_workspace(ctx)) {
    const std::string e = ctx->err;
    g3a_destroy(ctx);
    return fail(src, G3A_E_HIP, "sibling workspace: " + e);
  }
  *out = ctx;
  return G3A_OK;
}

// Workspace of each stage, grown to the batch's sizes (the same calls size it
// at set-up from max_reads, so that a first batch of typical shape finds it
// allocated).
// Grids of a call on n reads: the resident maxima for g3a_bwa_align; for the
// allocation-free call no more seeding threads than reads, and smaller
// traceback and extension grids, so that their per-thread pools fit the
// scratch array.
static void set_grids(g3a_ctx *ctx, int n, bool noalloc) {
  if (!ctx->seed_threads) ctx->seed_threads = seed_grid_threads(ctx->n_cu, ctx->idx, ctx->seed_block);
  if (!ctx->tb_threads) ctx->tb_threads = traceback_grid_threads(ctx->n_cu);
  ctx->run_seed_threads = ctx->seed_threads;
  ctx->run_tb_threads = ctx->tb_threads;
  ctx->run_lane_blocks = ctx->n_cu * ctx->lane_blocks_per_cu;
  ctx->run_slow_chunk = 16384;
  if (noalloc) {
    const int need = ((n + 127) / 128) * 128;
    ctx->run_seed_threads = std::max(ctx->seed_block, std::min(ctx->seed_threads, need));
    ctx->run_tb_threads = std::min(ctx->tb_threads, 4096);
    ctx->run_lane_blocks = std::min(ctx->run_lane_blocks, 64);
    ctx->run_slow_chunk = 16;
  }
}

static int alloc_fail(const g3a_ctx *ctx) { return ctx->arena ? G3A_E_ARG : G3A_E_HIP; }

static int alloc_reads(g3a_ctx *ctx, int n) {
  if (ensure(ctx, ctx->w->n_intv, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->n_seed, (size_t)n * 8 + 8) ||
      ensure(ctx, ctx->w->slow, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->intv_base, (size_t)n * 8 + 8) ||
      ensure(ctx, ctx->w->seed_off, (size_t)(n + 1) * 8) || ensure(ctx, ctx->w->counters, 64 * 8) ||
      ensure(ctx, ctx->w->n_task, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->task_off, (size_t)(n + 1) * 8) ||
      ensure(ctx, ctx->w->chain_big, (size_t)n * 4 + 4) ||
      ensure(ctx, ctx->w->n_aln, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->aln_off, (size_t)(n + 1) * 8) ||
      ensure(ctx, ctx->w->slow_list, (size_t)n * 4 + 4) ||
      ensure(ctx, ctx->w->scan_tmp, scan_tmp_bytes((ctx->arena ? n : ctx->max_reads) + 1)) ||
      ensure(ctx, ctx->w->intv, ((size_t)n * kSlotCap + 1024) * sizeof(Intv)))
    return G3A_E_HIP;
  if (ensure(ctx, ctx->w->seed_key, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->seed_key2, (size_t)n * 4 + 4) ||
      ensure(ctx, ctx->w->seed_idx, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->seed_order, (size_t)n * 4 + 4) ||
      ensure(ctx, ctx->w->seed_sort_tmp, seed_sort_tmp_bytes(n)))
    return G3A_E_HIP;
  return ensure(ctx, ctx->w->lep_pool, (size_t)ctx->run_seed_threads * (kLepCap + kFoundCap) * sizeof(Intv));
}

static int alloc_seeds(g3a_ctx *ctx, int64_t T_s, int n) {
  const size_t n_nodes = (size_t)((T_s + 32 * (int64_t)(n + 1)) / 4 + 2);
  return ensure(ctx, ctx->w->seeds, (size_t)T_s * sizeof(Seed)) || ensure(ctx, ctx->w->seeds2, (size_t)T_s * sizeof(Seed)) ||
         ensure(ctx, ctx->w->ch, (size_t)T_s * sizeof(ChainTmp)) || ensure(ctx, ctx->w->nxt, (size_t)T_s * 4) ||
         ensure(ctx, ctx->w->order, (size_t)T_s * 4) || ensure(ctx, ctx->w->tmpi, (size_t)T_s * 8 + 8) ||
         ensure(ctx, ctx->w->w16, (size_t)T_s * 2) || ensure(ctx, ctx->w->kept, (size_t)T_s) ||
         ensure(ctx, ctx->w->nodes, n_nodes * sizeof(KbNode)) || ensure(ctx, ctx->w->tasks_tmp, (size_t)T_s * sizeof(Task));
}

static int alloc_tasks(g3a_ctx *ctx, int64_t T_t) {
  return ensure(ctx, ctx->w->tasks, (size_t)T_t * sizeof(Task)) || ensure(ctx, ctx->w->regs, (size_t)T_t * sizeof(Region)) ||
         ensure(ctx, ctx->w->flag, (size_t)T_t) || ensure(ctx, ctx->w->sec, (size_t)T_t * 4);
}

static int alloc_alns(g3a_ctx *ctx, int64_t T_a) {
  return ensure(ctx, ctx->w->aln_reg, (size_t)T_a * sizeof(Region)) || ensure(ctx, ctx->w->aln_read, (size_t)T_a * 4) ||
         ensure(ctx, ctx->w->aln_rank, (size_t)T_a * 4) || ensure(ctx, ctx->w->z_size, (size_t)T_a * 8) ||
         ensure(ctx, ctx->w->cig_size, (size_t)T_a * 8) || ensure(ctx, ctx->w->z_off, (size_t)(T_a + 1) * 8) ||
         ensure(ctx, ctx->w->cig_off, (size_t)(T_a + 1) * 8) || ensure(ctx, ctx->w->alns, (size_t)T_a * sizeof(Aln)) ||
         ensure(ctx, ctx->w->scan_tmp, scan_tmp_bytes(T_a + 1)) || ensure(ctx, ctx->w->tb_order, (size_t)T_a * 4 + 4) ||
         ensure(ctx, ctx->w->tb_nw, (size_t)T_a * 4 + 4) ||
         ensure(ctx, ctx->w->tb_ctr, 64) || ensure(ctx, ctx->w->eh_pool, (size_t)traceback_pool_words(ctx->run_tb_threads) * 4);
}

static int alloc_paths(g3a_ctx *ctx, int64_t T_z, int64_t T_c) {
  return ensure(ctx, ctx->w->z, (size_t)T_z + 64) || ensure(ctx, ctx->w->cigar, (size_t)T_c * 4 + 64);
}

// Set-up sizes for a batch of max_reads reads and max_bases bases: 12 seeds,
// 4 extension tasks and one alignment per read, direction bytes for
// 4 x max_bases and the CIGAR capacity of one alignment per read.
static void set_grids(g3a_ctx *ctx, int n, bool noalloc);

static int reserve_workspace(g3a_ctx *ctx) {
  const int n = ctx->max_reads;
  set_grids(ctx, n, false);
  const int64_t nb = ctx->max_bases;
  if (alloc_reads(ctx, n) || alloc_seeds(ctx, 12 * (int64_t)n, n) || alloc_tasks(ctx, 4 * (int64_t)n) ||
      alloc_alns(ctx, n) || alloc_paths(ctx, 4 * nb, 4 * nb + 4 * (int64_t)n))
    return G3A_E_HIP;
  // host copy of the alignment records and CIGARs (32 words per read),
  // written once here so its pages are resident before the first batch
  if (ensure(ctx, ctx->w->cig_words, (size_t)n * 4 + 4) || ensure(ctx, ctx->w->cig_coff, (size_t)(n + 1) * 8) ||
      ensure(ctx, ctx->w->cig_out, (size_t)n * 128 + 64))
    return G3A_E_HIP;
  ctx->out_buf.assign((size_t)n * (sizeof(Aln) + 128) + 256, 0);
  return 0;
}

int g3a_memcpy_index(g3a_ctx *ctx) {
  if (!ctx) return G3A_E_ARG;
  if (ctx->store) return fail(ctx, G3A_E_STATE, "the device index is shared with sibling contexts");
  HIPCK(hipSetDevice(ctx->device));
  const HostIndex &h = ctx->h;
  DevIndex &x = ctx->idx;
  const uint64_t rows = h.seq_len + 1;
  if (rows > (4ULL << 32)) return fail(ctx, G3A_E_INDEX, "index larger than 2^34 rows");
  const bool narrow = rows < (1ULL << 32) && !ctx->force_wide;
  // Occ lines, built on the device from the BWA .bwt body (k_occ_lines)
  {
    DevBuf words;
    if (ensure(ctx, words, h.bwt_words.size() * 4)) return G3A_E_HIP;
    if (hipMemcpy(words.p, h.bwt_words.data(), h.bwt_words.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      free_buf(words);
      return fail(ctx, G3A_E_HIP, "copy of the .bwt body to the device failed");
    }
    const size_t wide_bytes = (size_t)((rows >> 6) + 1) * 32, narrow_bytes = (size_t)((rows >> 5) + 1) * 16,
                 sb_bytes = (size_t)((rows >> kNarrowSbShift) + 1) * 16;
    if (ensure(ctx, ctx->occ, wide_bytes)) { free_buf(words); return G3A_E_HIP; }
    if (narrow && (ensure(ctx, ctx->occn, narrow_bytes) || ensure(ctx, ctx->sbc, sb_bytes))) {
      free_buf(words);
      return G3A_E_HIP;
    }
    launch_occ_lines(P<uint32_t>(words), h.seq_len, h.primary, P<uint32_t>(ctx->occ),
                     narrow ? P<uint32_t>(ctx->occn) : nullptr, narrow ? P<uint32_t>(ctx->sbc) : nullptr,
                     ctx->stream);
    const hipError_t e1 = hipStreamSynchronize(ctx->stream), e2 = hipGetLastError();
    free_buf(words);
    if (e1 != hipSuccess || e2 != hipSuccess)
      return fail(ctx, G3A_E_HIP, std::string("Occ line build: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    // counts before each 2^32-row superblock of the wide lines
    for (uint64_t sb = 0; sb < 4; ++sb) {
      uint64_t c[4] = {0, 0, 0, 0};
      if ((sb << 32) < rows) bwa_occ_rows(h.bwt_words.data(), h.primary, sb << 32, c);
      for (int k = 0; k < 4; ++k) x.super[sb][k] = (int64_t)c[k];
    }
  }
  // 16 zero bytes of padding: word-wise readers may read past the last base
  if (ensure(ctx, ctx->pac, h.pac.size() + 16)) return G3A_E_HIP;
  HIPCK(hipMemset(ctx->pac.p, 0, h.pac.size() + 16));
  HIPCK(hipMemcpy(ctx->pac.p, h.pac.data(), h.pac.size(), hipMemcpyHostToDevice));
  if (ensure(ctx, ctx->sa, h.sa.size())) return G3A_E_HIP;
  HIPCK(hipMemcpy(ctx->sa.p, h.sa.data(), h.sa.size(), hipMemcpyHostToDevice));
  std::vector<int64_t> co, cl;
  std::vector<int32_t> ca;
  for (auto &c : h.anns) { co.push_back(c.offset); cl.push_back(c.len); ca.push_back(c.is_alt); }
  if (ensure(ctx, ctx->ctg_off, co.size() * 8) || ensure(ctx, ctx->ctg_len, cl.size() * 8) ||
      ensure(ctx, ctx->ctg_alt, ca.size() * 4))
    return G3A_E_HIP;
  HIPCK(hipMemcpy(ctx->ctg_off.p, co.data(), co.size() * 8, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(ctx->ctg_len.p, cl.data(), cl.size() * 8, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(ctx->ctg_alt.p, ca.data(), ca.size() * 4, hipMemcpyHostToDevice));
  {  // contig names and the MAPQ log table, for the record formatter
    std::string all;
    std::vector<int64_t> noff(1, 0);
    for (auto &c : h.anns) {
      all += c.name;
      noff.push_back((int64_t)all.size());
    }
    std::vector<float> lt(kLogTab);
    for (int i = 0; i < kLogTab; ++i) lt[(size_t)i] = logf((float)i);
    if (ensure(ctx, ctx->ctg_names, all.size() + 1) || ensure(ctx, ctx->ctg_name_off, noff.size() * 8) ||
        ensure(ctx, ctx->log_tab, lt.size() * 4))
      return G3A_E_HIP;
    HIPCK(hipMemcpy(ctx->ctg_names.p, all.c_str(), all.size() + 1, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->ctg_name_off.p, noff.data(), noff.size() * 8, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(ctx->log_tab.p, lt.data(), lt.size() * 4, hipMemcpyHostToDevice));
  }
  x.occ = P<uint32_t>(ctx->occ);
  build_counts(h, x.count);
  x.primary = (int64_t)h.primary;
  x.n = (int64_t)h.seq_len;
  x.l_pac = h.l_pac;
  x.pac = P<uint8_t>(ctx->pac);
  x.sa32 = h.sa_width == 4 ? P<uint32_t>(ctx->sa) : nullptr;
  x.sa64 = h.sa_width == 8 ? P<uint64_t>(ctx->sa) : nullptr;
  x.sa_shift = log2_exact(h.sa_intv);
  x.n_contigs = (int)h.anns.size();
  x.ctg_off = P<int64_t>(ctx->ctg_off);
  x.ctg_len = P<int64_t>(ctx->ctg_len);
  x.ctg_alt = P<int32_t>(ctx->ctg_alt);
  x.kmer = nullptr;
  x.occn = x.sbc = nullptr;
  if (narrow) {  // 32-bit interval fields
    x.occn = P<uint32_t>(ctx->occn);
    x.sbc = P<uint32_t>(ctx->sbc);
    if (ensure(ctx, ctx->kmer, ((size_t)3 << (2 * kKmer)) * 4)) return G3A_E_HIP;
    launch_kmer_table(x, P<uint32_t>(ctx->kmer), ctx->stream);
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipGetLastError());
    x.kmer = P<uint32_t>(ctx->kmer);
  }
  if (int rc = reserve_workspace(ctx)) return rc;
  ctx->index_on_device = true;
  return G3A_OK;
}

int g3a_memcpy_input(g3a_ctx *ctx, int n_reads, const uint8_t *seq, const int32_t *offsets) {
  if (!ctx || n_reads < 0 || (n_reads > 0 && (!seq || !offsets))) return G3A_E_ARG;
  if (n_reads > ctx->max_reads) return fail(ctx, G3A_E_ARG, "batch larger than max_reads");
  if (n_reads > 0) {
    if (offsets[0] != 0) return fail(ctx, G3A_E_ARG, "offsets[0] must be 0");
    for (int i = 0; i < n_reads; ++i)
      if (offsets[i + 1] < offsets[i]) return fail(ctx, G3A_E_ARG, "offsets must be non-decreasing");
    if (offsets[n_reads] > ctx->max_bases) return fail(ctx, G3A_E_ARG, "batch larger than max_bases");
  }
  HIPCK(hipSetDevice(ctx->device));
  const int64_t nb = n_reads ? offsets[n_reads] : 0;
  const size_t off_bytes = (size_t)(n_reads + 1) * 4;
  if (ensure(ctx, ctx->seq, (size_t)nb + 64) || ensure(ctx, ctx->off, off_bytes)) return G3A_E_HIP;
  if (int rc = stage_h2d(ctx, ctx->seq.p, seq, (size_t)nb)) return rc;
  if (n_reads)
    if (int rc = stage_h2d(ctx, ctx->off.p, offsets, off_bytes)) return rc;
  launch_clamp_codes(P<uint8_t>(ctx->seq), nb, ctx->stream);
  HIPCK(hipStreamSynchronize(ctx->stream));
  ctx->n_reads = n_reads;
  ctx->n_bases = nb;
  ctx->has_input = true;
  ctx->has_output = false;
  ctx->has_text = false;
  ctx->txt_bytes = -1;
  return G3A_OK;
}

int g3a_memcpy_text(g3a_ctx *ctx, int n_reads, const char *names, const int64_t *name_off, const char *quals,
                    const int64_t *qual_off) {
  if (!ctx || n_reads < 0 || (n_reads > 0 && (!names || !name_off || !quals || !qual_off))) return G3A_E_ARG;
  if (!ctx->has_input || n_reads != ctx->n_reads)
    return fail(ctx, G3A_E_STATE, "g3a_memcpy_text needs the batch of the last g3a_memcpy_input (same n_reads)");
  for (int i = 0; i < n_reads; ++i)
    if (name_off[i + 1] < name_off[i] || qual_off[i + 1] < qual_off[i])
      return fail(ctx, G3A_E_ARG, "name and quality offsets must be non-decreasing");
  if (n_reads > 0 && (name_off[0] != 0 || qual_off[0] != 0))
    return fail(ctx, G3A_E_ARG, "name_off[0] and qual_off[0] must be 0");
  HIPCK(hipSetDevice(ctx->device));
  const int64_t nb = n_reads ? name_off[n_reads] : 0, qb = n_reads ? qual_off[n_reads] : 0;
  const size_t ob = (size_t)(n_reads + 1) * 8;
  if (ensure(ctx, ctx->names, (size_t)nb + 16) || ensure(ctx, ctx->quals, (size_t)qb + 16) ||
      ensure(ctx, ctx->name_off, ob) || ensure(ctx, ctx->qual_off, ob))
    return G3A_E_HIP;
  if (n_reads > 0) {
    if (int rc = stage_h2d(ctx, ctx->names.p, names, (size_t)nb)) return rc;
    if (int rc = stage_h2d(ctx, ctx->quals.p, quals, (size_t)qb)) return rc;
    if (int rc = stage_h2d(ctx, ctx->name_off.p, name_off, ob)) return rc;
    if (int rc = stage_h2d(ctx, ctx->qual_off.p, qual_off, ob)) return rc;
  } else {
    HIPCK(hipMemsetAsync(ctx->name_off.p, 0, 8, ctx->stream));
    HIPCK(hipMemsetAsync(ctx->qual_off.p, 0, 8, ctx->stream));
    HIPCK(hipStreamSynchronize(ctx->stream));
  }
  ctx->has_text = true;
  ctx->text_n = n_reads;
  return G3A_OK;
}

namespace {

// read back k int64 values that sit in device memory at the given addresses
int readback(g3a_ctx *ctx, std::initializer_list<const int64_t *> src) {
  int i = 0;
  for (const int64_t *p : src) {
    HIPCK(hipMemcpyAsync(ctx->pinned + i, p, 8, hipMemcpyDeviceToHost, ctx->stream));
    ++i;
  }
  HIPCK(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // namespace


enum : int32_t {
  kOvSlow = 1,    // slow reads over caps.slow
  kOvPool = 2,    // their intervals over caps.pool
  kOvSeeds = 4,   // seeds
  kOvTasks = 8,   // extension tasks
  kOvAlns = 16,   // alignments
  kOvZ = 32,      // direction bytes
  kOvCig = 64,    // CIGAR words (traceback slots)
  kOvUsed = 128,  // CIGAR words used
};

// counters (int64 slots of w->counters): 0 FM steps, 8 slow-read count
// (int32), 16 chain error (int32), 24 extension cells, 32..37 learned-form
// counters, 40 overflow flag (int32), 41 k_seed's read counter (int32)
constexpr int kFlagSlot32 = 80;      // the flag, as an int32 index
constexpr int kNextReadSlot32 = 82;  // k_seed's read counter (int32)
constexpr int kChainBigSlot32 = 84;  // count of k_chain's large reads (int32)

static int64_t scale(double per, int64_t n, int64_t slack) { return (int64_t)(per * (double)n) + slack; }

static Caps plan_caps(const g3a_ctx *ctx, int n) {
  Caps c;
  c.seeds = scale(ctx->per_read[0], n, 4096);
  c.tasks = scale(ctx->per_read[1], n, 1024);
  c.alns = scale(ctx->per_read[2], n, 1024);
  c.used = scale(ctx->per_read[3], n, 1024);
  c.z = scale(ctx->per_base[0], ctx->n_bases, 65536);
  c.cig = scale(ctx->per_base[1], ctx->n_bases, 4 * (int64_t)n + 1024);
  const int slow_cap = ctx->arena ? 64 : ctx->slow_cap;  // the allocation-free path: few slow-path launches
  c.slow = n < slow_cap ? n : slow_cap;
  c.pool = (int64_t)c.slow * ctx->pool_per_slow + 1024;
  return c;
}

static void free_graph(g3a_ctx *ctx) {
  if (ctx->gexec) hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
  ctx->gkey.clear();
}

namespace {

// Buffers of the batched form's extension stage for `cap` tasks.
int alloc_ext(g3a_ctx *ctx, int64_t cap) {
  Work &w = *ctx->w;
  const int lane_blocks = ctx->arena ? ctx->run_lane_blocks : ctx->n_cu * ctx->lane_blocks_per_cu;
  return ensure(ctx, w.ext_res[0], (size_t)cap * sizeof(ExtRes)) ||
         ensure(ctx, w.ext_res[1], (size_t)cap * sizeof(ExtRes)) || ensure(ctx, w.ext_key, (size_t)cap * 4) ||
         ensure(ctx, w.ext_sorted, (size_t)cap * 4) || ensure(ctx, w.ext_slow, (size_t)cap * 4) ||
         ensure(ctx, w.ext_ctr, 2 * 4 * sizeof(int32_t)) || ensure(ctx, w.ext_hist, 1024 * 4) ||
         ensure(ctx, w.ext_rowbuf, (size_t)lane_blocks * kExtQCap * 64 * 4);
}

// Buffers of one attempt (allocation, if any, happens here and nowhere in
// enqueue_batch, so that the launches can be captured).
int alloc_batch(g3a_ctx *ctx, int n, const Caps &c) {
  Work &w = *ctx->w;
  if (alloc_reads(ctx, n)) return alloc_fail(ctx);
  const size_t fast_entries = (size_t)n * kSlotCap;
  const int chunk = ctx->run_slow_chunk < c.slow ? ctx->run_slow_chunk : c.slow;
  const int64_t scan_n = std::max<int64_t>({(int64_t)n, c.alns, (int64_t)c.slow}) + 1;
  if (ensure(ctx, w.intv, (fast_entries + (size_t)c.pool + 1024) * sizeof(Intv)) ||
      ensure(ctx, w.slow_scratch, (size_t)(chunk > 0 ? chunk : 1) * kSlowScratch * sizeof(Intv)) ||
      ensure(ctx, w.slow_off, (size_t)(c.slow + 1) * 8) || ensure(ctx, w.slow_n, (size_t)c.slow * 8 + 8) ||
      ensure(ctx, w.scan_tmp, scan_tmp_bytes(scan_n)) || ensure(ctx, w.totals, 16 * 8) ||
      alloc_seeds(ctx, c.seeds, n) || alloc_tasks(ctx, c.tasks) || alloc_alns(ctx, c.alns) ||
      alloc_paths(ctx, c.z, c.cig) || ensure(ctx, w.cig_words, (size_t)c.alns * 4 + 4) ||
      ensure(ctx, w.cig_coff, (size_t)(c.alns + 1) * 8) || ensure(ctx, w.cig_out, (size_t)c.used * 4 + 64))
    return alloc_fail(ctx);
  if (ctx->form != G3A_FORM_PLAIN && alloc_ext(ctx, c.tasks)) return alloc_fail(ctx);
  return 0;
}

}  // namespace

// The extension stage over the tasks in ctx->w->tasks (capacity cap, count
// at n_dev on the device), into ctx->w->regs.  Its buffers are allocated
// (alloc_batch) before it is called.
static int run_extension(g3a_ctx *ctx, int64_t cap, const int64_t *n_dev, int64_t *cells, bool timed) {
  hipStream_t s = ctx->stream;
  const int32_t *stop = P<int32_t>(ctx->w->counters) + kFlagSlot32;
  if (ctx->form == G3A_FORM_PLAIN) {
    ExtArgs ea{};
    ea.stop = stop;
    ea.idx = ctx->idx;
    ea.opt = ctx->opt;
    ea.n_tasks = cap;
    ea.n_dev = n_dev;
    ea.seq = P<uint8_t>(ctx->seq);
    ea.off = P<int32_t>(ctx->off);
    ea.tasks = P<Task>(ctx->w->tasks);
    ea.regs = P<Region>(ctx->w->regs);
    ea.cells = cells;
    if (timed) HIPCK(hipEventRecord(ctx->kev[3], s));
    launch_extend(ea, s);
    if (timed) HIPCK(hipEventRecord(ctx->kev[7], s));
    ctx->kev_used[3] = timed;
  } else if (cap > 0) {
    const int lane_blocks = ctx->arena ? ctx->run_lane_blocks : ctx->n_cu * ctx->lane_blocks_per_cu,
              wave_blocks = ctx->n_cu * ctx->wave_blocks_per_cu;
    launch_fill_i32(P<int32_t>(ctx->w->ext_ctr), 8, 0, s);
    for (int side = 0; side < 2; ++side) {
      ExtSideArgs ea{};
      ea.stop = stop;
      ea.idx = ctx->idx;
      ea.opt = ctx->opt;
      ea.side = side;
      ea.lane_opt_ok = ext_lane_opt_ok(ctx->opt);
      ea.n_tasks = cap;
      ea.n_dev = n_dev;
      ea.seq = P<uint8_t>(ctx->seq);
      ea.off = P<int32_t>(ctx->off);
      ea.tasks = P<Task>(ctx->w->tasks);
      ea.left = P<ExtRes>(ctx->w->ext_res[0]);
      ea.res = P<ExtRes>(ctx->w->ext_res[side]);
      ea.key = P<uint32_t>(ctx->w->ext_key);
      ea.sorted = P<int32_t>(ctx->w->ext_sorted);
      ea.slow_list = P<int32_t>(ctx->w->ext_slow);
      ea.ctr = P<int32_t>(ctx->w->ext_ctr) + 4 * side;
      ea.rowbuf = P<uint32_t>(ctx->w->ext_rowbuf);
      ea.cells = cells;
      launch_ext_classify(ea, s);
      ext_sort(P<uint32_t>(ctx->w->ext_key), cap, n_dev, P<int32_t>(ctx->w->ext_hist), P<int32_t>(ctx->w->ext_sorted), s);
      if (timed) HIPCK(hipEventRecord(ctx->kev[2 + 2 * side], s));
      launch_ext_lane(ea, ctx->strip_rows, ctx->ext_packed, lane_blocks, s);
      if (timed) HIPCK(hipEventRecord(ctx->kev[3 + 2 * side], s));
      launch_ext_wave(ea, wave_blocks, s);
      if (timed) HIPCK(hipEventRecord(ctx->kev[7 + 2 * side], s));
      ctx->kev_used[1 + side] = timed;
    }
    ExtFinishArgs fa{};
    fa.stop = stop;
    fa.n_tasks = cap;
    fa.n_dev = n_dev;
    fa.tasks = P<Task>(ctx->w->tasks);
    fa.left = P<ExtRes>(ctx->w->ext_res[0]);
    fa.right = P<ExtRes>(ctx->w->ext_res[1]);
    fa.regs = P<Region>(ctx->w->regs);
    launch_ext_finish(fa, s);
  }
  return 0;
}

// Slots of the totals read back after a batch (w->totals).
enum { kTotSlow, kTotPool, kTotSeeds, kTotTasks, kTotAlns, kTotZ, kTotCig, kTotUsed, kTotChainErr, kTotFlags, kTotN };

// Every launch of one attempt at the resident batch, on ctx->stream, with
// no host synchronization and no allocation: the sequence a graph captures.
// timed: record the stage and kernel events (not while capturing).
static int enqueue_batch(g3a_ctx *ctx, int n, const Caps &c, bool timed) {
  hipStream_t s = ctx->stream;
  Work &w = *ctx->w;
  int64_t *cnt = P<int64_t>(w.counters);
  int32_t *flag = P<int32_t>(w.counters) + kFlagSlot32;
  int32_t *n_slow_dev = P<int32_t>(w.counters) + 2 * 8;
  const size_t fast_entries = (size_t)n * kSlotCap;
  launch_fill_i32(P<int32_t>(w.counters), 128, 0, s);  // a kernel, not a memset node
  if (timed) HIPCK(hipEventRecord(ctx->ev[0], s));
  // ---- seeding
  SeedArgs sa{};
  sa.stop = flag;
  sa.idx = ctx->idx;
  sa.opt = ctx->opt;
  sa.n_reads = n;
  sa.seq = P<uint8_t>(ctx->seq);
  sa.off = P<int32_t>(ctx->off);
  sa.intv = P<Intv>(w.intv);
  sa.slot_cap = kSlotCap;
  sa.n_intv = P<int32_t>(w.n_intv);
  sa.n_seed = P<int64_t>(w.n_seed);
  sa.slow = P<int32_t>(w.slow);
  sa.fm_steps = cnt + 0;
  sa.lep_pool = P<Intv>(w.lep_pool);
  sa.li_stats = ctx->idx.saf ? cnt + 32 : nullptr;
  sa.next_read = P<int32_t>(w.counters) + kNextReadSlot32;
  sa.order = nullptr;
  if (ctx->seed_ordered && n > 1) {
    launch_seed_keys(sa.seq, sa.off, n, P<uint32_t>(w.seed_key), P<int32_t>(w.seed_idx), s);
    if (seed_sort(P<uint32_t>(w.seed_key), P<uint32_t>(w.seed_key2), P<int32_t>(w.seed_idx),
                  P<int32_t>(w.seed_order), n, w.seed_sort_tmp.p, w.seed_sort_tmp.bytes, s))
      return fail(ctx, G3A_E_HIP, "seed order sort failed");
    sa.order = P<int32_t>(w.seed_order);
  }
  if (timed) HIPCK(hipEventRecord(ctx->kev[0], s));
  const int seed_run =
      ctx->seed_blocks_per_cu > 0 && ctx->seed_blocks_per_cu * ctx->seed_block * ctx->n_cu < ctx->run_seed_threads
          ? ctx->seed_blocks_per_cu * ctx->seed_block * ctx->n_cu
          : ctx->run_seed_threads;
  launch_seed(sa, seed_run, ctx->seed_blocks_per_cu, ctx->seed_block, s);
  if (timed) HIPCK(hipEventRecord(ctx->kev[1], s));
  ctx->kev_used[0] = timed;
  launch_collect_slow(n, P<int32_t>(w.slow), P<int32_t>(w.slow_list), n_slow_dev, P<int64_t>(w.intv_base), kSlotCap, s);
  // reads whose lists outgrew the fast path: counts (pass 0), pool offsets,
  // then the intervals (pass 1), all for the device's count of them
  launch_cap_check32(n_slow_dev, c.slow, flag, kOvSlow, s);
  {
    const int chunk = ctx->run_slow_chunk < c.slow ? ctx->run_slow_chunk : c.slow;
    const int32_t *list = P<int32_t>(w.slow_list);
    for (int c0 = 0; c0 < c.slow; c0 += chunk) {
      const int m = c.slow - c0 < chunk ? c.slow - c0 : chunk;
      launch_seed_slow(sa, list + c0, m, n_slow_dev, c0, P<Intv>(w.slow_scratch), nullptr, nullptr, 0, 0, s);
    }
    launch_gather_slow_counts(c.slow, n_slow_dev, list, P<int32_t>(w.n_intv), P<int64_t>(w.slow_n), s);
    scan_i64(P<int64_t>(w.slow_n), P<int64_t>(w.slow_off), c.slow, w.scan_tmp.p, w.scan_tmp.bytes, s);
    launch_cap_check(P<int64_t>(w.slow_off) + c.slow, c.pool, flag, kOvPool, s);
    Intv *pool_base = P<Intv>(w.intv) + fast_entries;
    for (int c0 = 0; c0 < c.slow; c0 += chunk) {
      const int m = c.slow - c0 < chunk ? c.slow - c0 : chunk;
      launch_seed_slow(sa, list + c0, m, n_slow_dev, c0, P<Intv>(w.slow_scratch), P<int64_t>(w.slow_off) + c0,
                       pool_base, 1, c.slow <= chunk, s);
    }
    launch_set_slow_base(c.slow, n_slow_dev, list, P<int64_t>(w.slow_off), (int64_t)fast_entries,
                         P<int64_t>(w.intv_base), s);
  }
  if (timed) HIPCK(hipEventRecord(ctx->ev[1], s));
  // ---- SA lookup (the seeds are sorted by the chaining kernel)
  scan_i64(P<int64_t>(w.n_seed), P<int64_t>(w.seed_off), n, w.scan_tmp.p, w.scan_tmp.bytes, s);
  const int64_t *T_s = P<int64_t>(w.seed_off) + n;
  launch_cap_check(T_s, c.seeds, flag, kOvSeeds, s);
  SaArgs la{};
  la.stop = flag;
  la.idx = ctx->idx;
  la.opt = ctx->opt;
  la.n_reads = n;
  la.total = c.seeds;
  la.total_dev = T_s;
  la.seed_off = P<int64_t>(w.seed_off);
  la.intv = P<Intv>(w.intv);
  la.intv_base = P<int64_t>(w.intv_base);
  la.n_intv = P<int32_t>(w.n_intv);
  la.seeds = P<Seed>(w.seeds);
  launch_sa_lookup(la, s);
  if (timed) HIPCK(hipEventRecord(ctx->ev[2], s));
  // ---- chaining
  ChainArgs ca{};
  ca.stop = flag;
  ca.idx = ctx->idx;
  ca.opt = ctx->opt;
  ca.n_reads = n;
  ca.off = P<int32_t>(ctx->off);
  ca.seed_off = P<int64_t>(w.seed_off);
  ca.seeds = P<Seed>(w.seeds);
  ca.sort_tmp = P<Seed>(w.seeds2);
  ca.ch = P<ChainTmp>(w.ch);
  ca.nxt = P<int32_t>(w.nxt);
  ca.order = P<int32_t>(w.order);
  ca.tmpi = P<int32_t>(w.tmpi);
  ca.w16 = P<uint16_t>(w.w16);
  ca.kept = P<uint8_t>(w.kept);
  ca.nodes = P<KbNode>(w.nodes);
  ca.tasks_tmp = P<Task>(w.tasks_tmp);
  ca.n_task = P<int32_t>(w.n_task);
  ca.error = P<int32_t>(w.counters) + 2 * 16;
  ca.serial = ctx->chain_serial;
  ca.big = P<int32_t>(w.chain_big);
  ca.n_big = P<int32_t>(w.counters) + kChainBigSlot32;
  launch_chain(ca, s);
  scan_i32(P<int32_t>(w.n_task), P<int64_t>(w.task_off), n, w.scan_tmp.p, w.scan_tmp.bytes, s);
  const int64_t *T_t = P<int64_t>(w.task_off) + n;
  launch_cap_check(T_t, c.tasks, flag, kOvTasks, s);
  CompactArgs cp{flag, n, P<int64_t>(w.seed_off), P<int64_t>(w.task_off), P<Task>(w.tasks_tmp), P<Task>(w.tasks)};
  launch_task_compact(cp, s);
  if (timed) HIPCK(hipEventRecord(ctx->ev[3], s));
  // ---- extension
  if (int rc = run_extension(ctx, c.tasks, T_t, cnt + 24, timed)) return rc;
  if (timed) HIPCK(hipEventRecord(ctx->ev[4], s));
  // ---- region filter, traceback, finalize
  RegionArgs ra{flag, ctx->opt, n, P<int64_t>(w.task_off), P<Region>(w.regs), P<uint8_t>(w.flag),
                P<int32_t>(w.sec), P<int32_t>(w.n_aln)};
  launch_region_filter(ra, s);
  scan_i32(P<int32_t>(w.n_aln), P<int64_t>(w.aln_off), n, w.scan_tmp.p, w.scan_tmp.bytes, s);
  const int64_t *T_a = P<int64_t>(w.aln_off) + n;
  launch_cap_check(T_a, c.alns, flag, kOvAlns, s);
  AlnPrepArgs ap{flag, ctx->opt, n, P<int64_t>(w.task_off), P<int64_t>(w.aln_off), P<Region>(w.regs),
                 P<Region>(w.aln_reg), P<int32_t>(w.aln_read), P<int32_t>(w.aln_rank),
                 P<int64_t>(w.z_size), P<int64_t>(w.cig_size)};
  launch_aln_prep(ap, s);
  // sizes past the alignments count nothing in the scans over the capacity
  launch_zero_tail(P<int64_t>(w.z_size), c.alns, T_a, s);
  launch_zero_tail(P<int64_t>(w.cig_size), c.alns, T_a, s);
  scan_i64(P<int64_t>(w.z_size), P<int64_t>(w.z_off), c.alns, w.scan_tmp.p, w.scan_tmp.bytes, s);
  scan_i64(P<int64_t>(w.cig_size), P<int64_t>(w.cig_off), c.alns, w.scan_tmp.p, w.scan_tmp.bytes, s);
  launch_cap_check(P<int64_t>(w.z_off) + c.alns, c.z, flag, kOvZ, s);
  launch_cap_check(P<int64_t>(w.cig_off) + c.alns, c.cig, flag, kOvCig, s);
  TbArgs tb{};
  tb.stop = flag;
  tb.idx = ctx->idx;
  tb.opt = ctx->opt;
  tb.n_alns = c.alns;
  tb.n_dev = T_a;
  tb.seq = P<uint8_t>(ctx->seq);
  tb.off = P<int32_t>(ctx->off);
  tb.aln_reg = P<Region>(w.aln_reg);
  tb.aln_read = P<int32_t>(w.aln_read);
  tb.aln_rank = P<int32_t>(w.aln_rank);
  tb.z_off = P<int64_t>(w.z_off);
  tb.cig_off = P<int64_t>(w.cig_off);
  tb.z = P<uint8_t>(w.z);
  tb.cigar = P<uint32_t>(w.cigar);
  tb.alns = P<Aln>(w.alns);
  tb.eh_pool = P<int32_t>(w.eh_pool);
  tb.tb_order = P<int32_t>(w.tb_order);
  tb.tb_ctr = P<int32_t>(w.tb_ctr);
  tb.tb_nw = P<int32_t>(w.tb_nw);
  if (timed) HIPCK(hipEventRecord(ctx->kev[10], s));
  launch_traceback(tb, ctx->run_tb_threads, ctx->n_cu, s);
  if (timed) HIPCK(hipEventRecord(ctx->kev[11], s));
  ctx->kev_used[5] = timed;
  // CIGARs packed to the words used, for the copy to the host
  launch_cig_words(P<Aln>(w.alns), c.alns, T_a, P<int32_t>(w.cig_words), s);
  scan_i32(P<int32_t>(w.cig_words), P<int64_t>(w.cig_coff), c.alns, w.scan_tmp.p, w.scan_tmp.bytes, s);
  launch_cap_check(P<int64_t>(w.cig_coff) + c.alns, c.used, flag, kOvUsed, s);
  launch_cig_gather(P<Aln>(w.alns), c.alns, T_a, flag, P<int64_t>(w.cig_coff), P<uint32_t>(w.cigar),
                    P<uint32_t>(w.cig_out), s);
  // the totals, for the one read-back of the batch
  Gather16 g{};
  g.src[kTotSlow] = cnt + 8;
  g.src[kTotPool] = P<int64_t>(w.slow_off) + c.slow;
  g.src[kTotSeeds] = T_s;
  g.src[kTotTasks] = T_t;
  g.src[kTotAlns] = T_a;
  g.src[kTotZ] = P<int64_t>(w.z_off) + c.alns;
  g.src[kTotCig] = P<int64_t>(w.cig_off) + c.alns;
  g.src[kTotUsed] = P<int64_t>(w.cig_coff) + c.alns;
  g.src[kTotChainErr] = cnt + 16;
  g.src[kTotFlags] = cnt + kFlagSlot32 / 2;
  launch_gather_totals(g, kTotN, P<int64_t>(w.totals), s);
  if (timed) HIPCK(hipEventRecord(ctx->ev[5], s));
  return 0;
}

// A byte string naming everything a captured batch bakes into its launches:
// a graph is replayed only while it is unchanged.
static std::string graph_key(const g3a_ctx *ctx, int n, const Caps &c) {
  std::string k;
  auto put = [&](const void *p, size_t b) { k.append((const char *)p, b); };
  put(&n, sizeof n);
  put(&c, sizeof c);
  put(&ctx->idx, sizeof ctx->idx);
  put(&ctx->opt, sizeof ctx->opt);
  const int v[] = {ctx->form, ctx->strip_rows, ctx->lane_blocks_per_cu, ctx->wave_blocks_per_cu, ctx->ext_packed,
                   ctx->seed_blocks_per_cu, ctx->seed_block, ctx->run_seed_threads, ctx->run_tb_threads, ctx->run_lane_blocks,
                   ctx->run_slow_chunk, ctx->chain_serial};
  put(v, sizeof v);
  put(&ctx->buf_gen, sizeof ctx->buf_gen);
  return k;
}

// The pipeline on the resident batch, with ctx->w as its workspace: one
// attempt per capacity plan, redone with the overflowing capacity grown
// until every count fits (the first attempt fits once the context has seen
// batches like this one).
static int align_batch(g3a_ctx *ctx) {
  hipStream_t s = ctx->stream;
  const int n = ctx->n_reads;
  ctx->has_output = false;
  ctx->total_seeds = ctx->total_tasks = ctx->total_alns = ctx->total_cig = ctx->n_slow = 0;
  Caps c = plan_caps(ctx, n);
  for (int attempt = 0;; ++attempt) {
    if (attempt == 8) return fail(ctx, G3A_E_INTERNAL, "batch capacities did not settle");
    if (ctx->arena) {  // an allocation-free attempt carves the scratch afresh
      ctx->arena->used = 0;
      ctx->scratch_work = Work{};
    }
    for (int &u : ctx->kev_used) u = 0;
    if (alloc_batch(ctx, n, c)) return alloc_fail(ctx);
    const bool graph = ctx->use_graph && !ctx->arena && !ctx->seed_ordered;
    ctx->graph_used = false;
    if (graph) {
      const std::string key = graph_key(ctx, n, c);
      if (!ctx->gexec || key != ctx->gkey) {
        free_graph(ctx);
        HIPCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batch(ctx, n, c, false);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &gr);
        if (rc) {
          if (gr) hipGraphDestroy(gr);
          return rc;
        }
        HIPCK(e);
        ctx->graph = gr;
        HIPCK(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
        ctx->gkey = key;
        ++ctx->graph_captures;
      }
      HIPCK(hipEventRecord(ctx->ev[0], s));
      HIPCK(hipGraphLaunch(ctx->gexec, s));
      HIPCK(hipEventRecord(ctx->ev[5], s));
      ctx->graph_used = true;
    } else {
      if (int rc = enqueue_batch(ctx, n, c, true)) return rc;
    }
    HIPCK(hipMemcpyAsync(ctx->pinned, P<int64_t>(ctx->w->totals), kTotN * 8, hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    HIPCK(hipGetLastError());
    int64_t t[kTotN];
    std::memcpy(t, ctx->pinned, sizeof t);
    const int32_t flags = (int32_t)(t[kTotFlags] & 0xffffffff);
    const int64_t n_slow = t[kTotSlow] & 0xffffffff;
    if (flags) {
      // the first count that did not fit (the kernels after it stopped, and
      // the checks after it read nothing): grow it past the count
      auto grow = [](int64_t have, int64_t need) { return std::max(2 * have, need + need / 4 + 1024); };
      if (flags & kOvSlow) c.slow = (int)std::min<int64_t>(n, grow(c.slow, n_slow));
      else if (flags & kOvPool) c.pool = grow(c.pool, t[kTotPool]);
      else if (flags & kOvSeeds) c.seeds = grow(c.seeds, t[kTotSeeds]);
      else if (flags & kOvTasks) c.tasks = grow(c.tasks, t[kTotTasks]);
      else if (flags & kOvAlns) c.alns = grow(c.alns, t[kTotAlns]);
      else {
        if (flags & kOvZ) c.z = grow(c.z, t[kTotZ]);
        if (flags & kOvCig) c.cig = grow(c.cig, t[kTotCig]);
        if (flags & kOvUsed) c.used = grow(c.used, t[kTotUsed]);
      }
      ++ctx->overflow_reruns;
      continue;
    }
    if (t[kTotChainErr] & 0xffffffff) return fail(ctx, G3A_E_INTERNAL, "chain node arena overflow");
    ctx->caps = c;
    ctx->n_slow = n_slow;
    ctx->total_seeds = t[kTotSeeds];
    ctx->total_tasks = t[kTotTasks];
    ctx->total_alns = t[kTotAlns];
    ctx->total_cig = t[kTotUsed];
    // next batches start from what this one needed, with headroom (buffers
    // only grow, so a smaller plan costs nothing but a new graph capture)
    if (n >= 1024 && !ctx->arena) {
      auto per = [&](double &r, int64_t total, int64_t units) {
        r = 1.25 * (double)total / (double)std::max<int64_t>(units, 1);
      };
      per(ctx->per_read[0], t[kTotSeeds], n);
      per(ctx->per_read[1], t[kTotTasks], n);
      per(ctx->per_read[2], t[kTotAlns], n);
      per(ctx->per_read[3], t[kTotUsed], n);
      per(ctx->per_base[0], t[kTotZ], ctx->n_bases);
      per(ctx->per_base[1], t[kTotCig], ctx->n_bases);
      if (n_slow > 0) ctx->pool_per_slow = std::max<int64_t>(ctx->pool_per_slow, 2 * t[kTotPool] / n_slow + 1);
      ctx->slow_cap = std::max<int>(ctx->slow_cap, (int)(n_slow + n_slow / 4) + 16);
    }
    break;
  }
  float tm[5];
  for (float &m : ctx->ms) m = 0;
  if (!ctx->graph_used) {
    for (int i = 0; i < 5; ++i) hipEventElapsedTime(&tm[i], ctx->ev[i], ctx->ev[i + 1]);
    for (int i = 0; i < 5; ++i) ctx->ms[i] = tm[i];
  }
  hipEventElapsedTime(&ctx->ms[5], ctx->ev[0], ctx->ev[5]);
  float e = 0;
  for (float &k : ctx->kms) k = 0;
  if (ctx->kev_used[0] && hipEventElapsedTime(&e, ctx->kev[0], ctx->kev[1]) == hipSuccess) ctx->kms[0] = e;
  for (int side = 0; side < 2; ++side)
    if (ctx->kev_used[1 + side]) {
      if (hipEventElapsedTime(&e, ctx->kev[2 + 2 * side], ctx->kev[3 + 2 * side]) == hipSuccess) ctx->kms[1] += e;
      if (hipEventElapsedTime(&e, ctx->kev[3 + 2 * side], ctx->kev[7 + 2 * side]) == hipSuccess) ctx->kms[2] += e;
    }
  if (ctx->kev_used[3] && hipEventElapsedTime(&e, ctx->kev[3], ctx->kev[7]) == hipSuccess) ctx->kms[2] = e;
  if (ctx->kev_used[5] && hipEventElapsedTime(&e, ctx->kev[10], ctx->kev[11]) == hipSuccess) ctx->kms[3] = e;
  ctx->has_output = true;
  return G3A_OK;
}

int g3a_bwa_align(g3a_ctx *ctx) {
  if (!ctx) return G3A_E_ARG;
  if (!ctx->index_on_device || !ctx->has_input) return fail(ctx, G3A_E_STATE, "index or input not on device");
  HIPCK(hipSetDevice(ctx->device));
  ctx->w = &ctx->own;
  set_grids(ctx, ctx->n_reads, false);
  return align_batch(ctx);
}

int g3a_bwa_align_noalloc(g3a_ctx *ctx) {
  if (!ctx) return G3A_E_ARG;
  if (!ctx->index_on_device || !ctx->has_input) return fail(ctx, G3A_E_STATE, "index or input not on device");
  if (ctx->n_reads > G3A_NOALLOC_MAX_READS)
    return fail(ctx, G3A_E_ARG, "the allocation-free entry point takes at most " +
                                    std::to_string(G3A_NOALLOC_MAX_READS) + " reads per batch");
  if (!ctx->scratch) return fail(ctx, G3A_E_STATE, "no scratch array on this device");
  HIPCK(hipSetDevice(ctx->device));
  // the previous contents of the scratch array are not read: every buffer
  // the pipeline expects to start at zero is cleared by a kernel
  ScratchArena arena;
  arena.base = ctx->scratch;
  arena.cap = (size_t)G3A_NOALLOC_SCRATCH_BYTES;
  ctx->scratch_work = Work{};
  ctx->w = &ctx->scratch_work;
  ctx->arena = &arena;
  set_grids(ctx, ctx->n_reads, true);
  const int rc = align_batch(ctx);
  ctx->arena = nullptr;
  if (rc) ctx->has_output = false;
  return rc;
}

static TextArgs text_args(g3a_ctx *ctx, int mode) {
  TextArgs t{};
  t.mode = mode;
  t.n_reads = ctx->n_reads;
  t.seq = P<uint8_t>(ctx->seq);
  t.off = P<int32_t>(ctx->off);
  t.names = P<char>(ctx->names);
  t.name_off = P<int64_t>(ctx->name_off);
  t.quals = P<char>(ctx->quals);
  t.qual_off = P<int64_t>(ctx->qual_off);
  t.ctg_names = P<char>(ctx->ctg_names);
  t.ctg_name_off = P<int64_t>(ctx->ctg_name_off);
  t.ctg_off = P<int64_t>(ctx->ctg_off);
  t.pac = ctx->idx.pac;
  t.l_pac = ctx->idx.l_pac;
  t.aln_off = P<int64_t>(ctx->w->aln_off);
  t.alns = P<Aln>(ctx->w->alns);
  t.cigar = P<uint32_t>(ctx->w->cig_out);  // the compacted CIGARs (k_cig_gather)
  t.log_tab = P<float>(ctx->log_tab);
  t.a = ctx->opt.a;
  t.b = ctx->opt.b;
  t.min_seed_len = ctx->opt.min_seed_len;
  return t;
}

int g3a_format_records(g3a_ctx *ctx, int mode, int64_t *bytes) {
  if (!ctx || (mode != 0 && mode != 1)) return G3A_E_ARG;
  if (!ctx->has_output) return fail(ctx, G3A_E_STATE, "no alignment results");
  if (!ctx->has_text || ctx->text_n != ctx->n_reads)
    return fail(ctx, G3A_E_STATE, "no names and qualities for this batch (g3a_memcpy_text)");
  HIPCK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int n = ctx->n_reads;
  if (ensure(ctx, ctx->txt_len, (size_t)n * 8 + 8) || ensure(ctx, ctx->txt_off, (size_t)(n + 1) * 8) ||
      ensure(ctx, ctx->fmt_scan_tmp, scan_tmp_bytes(n + 1)))
    return G3A_E_HIP;
  const TextArgs t = text_args(ctx, mode);
  // one formatting pass into a staging buffer laid out by per-read bounds,
  // then the texts moved to their places; should a read outgrow its bound,
  // the texts are made again by a counting pass and a writing pass
  if (ensure(ctx, ctx->txt_bound, (size_t)n * 8 + 8) || ensure(ctx, ctx->stage_off, (size_t)(n + 1) * 8) ||
      ensure(ctx, ctx->fmt_flag, 8))
    return G3A_E_HIP;
  launch_text_bound(t, P<int64_t>(ctx->txt_bound), s);
  scan_i64(P<int64_t>(ctx->txt_bound), P<int64_t>(ctx->stage_off), n, ctx->fmt_scan_tmp.p, ctx->fmt_scan_tmp.bytes,
           s);
  if (readback(ctx, {P<int64_t>(ctx->stage_off) + n})) return G3A_E_HIP;
  if (ensure(ctx, ctx->txt_stage, (size_t)ctx->pinned[0] + 16)) return G3A_E_HIP;
  launch_fill_i32(P<int32_t>(ctx->fmt_flag), 2, 0, s);
  launch_text_staged(t, P<int64_t>(ctx->txt_len), P<int64_t>(ctx->stage_off), P<char>(ctx->txt_stage), s);
  scan_i64(P<int64_t>(ctx->txt_len), P<int64_t>(ctx->txt_off), n, ctx->fmt_scan_tmp.p, ctx->fmt_scan_tmp.bytes, s);
  launch_text_check(n, P<int64_t>(ctx->txt_len), P<int64_t>(ctx->txt_bound), P<int64_t>(ctx->fmt_flag), s);
  if (readback(ctx, {P<int64_t>(ctx->txt_off) + n, P<int64_t>(ctx->fmt_flag)})) return G3A_E_HIP;
  int64_t total = ctx->pinned[0];
  if (ctx->pinned[1]) {
    launch_text(t, P<int64_t>(ctx->txt_len), nullptr, nullptr, false, s);
    scan_i64(P<int64_t>(ctx->txt_len), P<int64_t>(ctx->txt_off), n, ctx->fmt_scan_tmp.p, ctx->fmt_scan_tmp.bytes, s);
    if (readback(ctx, {P<int64_t>(ctx->txt_off) + n})) return G3A_E_HIP;
    total = ctx->pinned[0];
    if (ensure(ctx, ctx->txt, (size_t)total + 16)) return G3A_E_HIP;
    launch_text(t, nullptr, P<int64_t>(ctx->txt_off), P<char>(ctx->txt), true, s);
  } else {
    if (ensure(ctx, ctx->txt, (size_t)total + 16)) return G3A_E_HIP;
    launch_text_pack(n, P<char>(ctx->txt_stage), P<int64_t>(ctx->stage_off), P<int64_t>(ctx->txt_len),
                     P<int64_t>(ctx->txt_off), P<char>(ctx->txt), s);
  }
  HIPCK(hipStreamSynchronize(s));
  HIPCK(hipGetLastError());
  ctx->txt_bytes = total;
  if (bytes) *bytes = total;
  return G3A_OK;
}

int g3a_copy_text(g3a_ctx *ctx, char *dst, int64_t cap) {
  if (!ctx || (!dst && cap > 0)) return G3A_E_ARG;
  if (ctx->txt_bytes < 0) return fail(ctx, G3A_E_STATE, "no formatted records (g3a_format_records)");
  if (cap < ctx->txt_bytes) return fail(ctx, G3A_E_ARG, "text buffer smaller than the formatted records");
  HIPCK(hipSetDevice(ctx->device));
  if (ctx->txt_bytes > 0)
    if (int rc = stage_d2h(ctx, dst, ctx->txt.p, (size_t)ctx->txt_bytes)) return rc;
  return G3A_OK;
}

static_assert(sizeof(g3a_pestat) == sizeof(PeStat) && offsetof(g3a_pestat, avg) == offsetof(PeStat, avg),
              "g3a_pestat and PeStat share one layout");

// Paired-end arguments over the last batch's alignments (reads 2p, 2p+1 are
// pair p).
static int pair_args(g3a_ctx *ctx, PairArgs &a) {
  if (!ctx->has_output) return fail(ctx, G3A_E_STATE, "no alignment results");
  if (ctx->n_reads & 1) return fail(ctx, G3A_E_ARG, "a paired batch holds an even number of reads");
  a = PairArgs{};
  a.n_pairs = ctx->n_reads / 2;
  a.aln_off = P<int64_t>(ctx->w->aln_off);
  a.alns = P<Aln>(ctx->w->alns);
  a.ctg_off = P<int64_t>(ctx->ctg_off);
  a.l_pac = ctx->idx.l_pac;
  const Opt &o = ctx->opt;
  a.a = o.a, a.b = o.b, a.o_del = o.o_del, a.e_del = o.e_del, a.o_ins = o.o_ins, a.e_ins = o.e_ins;
  a.min_seed_len = o.min_seed_len;
  a.mask_level = o.mask_level;
  return 0;
}

int g3a_mem_pestat(g3a_ctx *ctx, int max_ins, g3a_pestat pes[4]) {
  if (!ctx || !pes) return G3A_E_ARG;
  if (max_ins < 0 || max_ins > G3A_PE_MAX_INS)
    return fail(ctx, G3A_E_ARG, "max_ins must lie in [0, " + std::to_string(G3A_PE_MAX_INS) + "]");
  HIPCK(hipSetDevice(ctx->device));
  PairArgs a;
  if (int rc = pair_args(ctx, a)) return rc;
  a.max_ins = max_ins;
  // up to 16 histogram copies, within 64 MiB
  a.copies = (int)std::max<size_t>(1, std::min<size_t>(16, ((size_t)64 << 20) / (16 * ((size_t)max_ins + 1))));
  if (int rc = ensure(ctx, ctx->pe_hist, 16 * ((size_t)max_ins + 1) * (size_t)a.copies)) return rc;
  if (int rc = ensure(ctx, ctx->pe_stat, 4 * sizeof(PeStat))) return rc;
  a.hist = P<uint32_t>(ctx->pe_hist);
  a.stat = P<PeStat>(ctx->pe_stat);
  launch_pe_stat(a, ctx->stream);
  HIPCK(hipGetLastError());
  return stage_d2h(ctx, pes, ctx->pe_stat.p, 4 * sizeof(PeStat));
}

int g3a_mem_pair(g3a_ctx *ctx, const g3a_pestat pes[4], int64_t id0, int32_t *out) {
  if (!ctx || !pes || !out) return G3A_E_ARG;
  HIPCK(hipSetDevice(ctx->device));
  PairArgs a;
  if (int rc = pair_args(ctx, a)) return rc;
  if (a.n_pairs == 0) return 0;
  if (int rc = ensure(ctx, ctx->pe_v, 16 * (size_t)std::max<int64_t>(ctx->total_alns, 1))) return rc;
  if (int rc = ensure(ctx, ctx->pe_out, 24 * (size_t)a.n_pairs)) return rc;
  std::memcpy(a.pes, pes, sizeof a.pes);
  a.v = P<uint64_t>(ctx->pe_v);
  a.out = P<int32_t>(ctx->pe_out);
  a.id0 = id0;
  launch_pe_pair(a, ctx->stream);
  HIPCK(hipGetLastError());
  return stage_d2h(ctx, out, ctx->pe_out.p, 24 * (size_t)a.n_pairs);
}

int g3a_memcpy_output(g3a_ctx *ctx, g3a_batch_out *out) {
  if (!ctx || !out) return G3A_E_ARG;
  if (!ctx->has_output) return fail(ctx, G3A_E_STATE, "no alignment results");
  HIPCK(hipSetDevice(ctx->device));
  std::memset(out, 0, sizeof *out);
  const int n = ctx->n_reads;
  const int64_t na = ctx->total_alns;
  out->n_reads = n;
  out->n_alns = na;
  out->aln_off = (int64_t *)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  // through the staging slots into a host buffer kept by the context, then
  // unpacked into the caller's arrays
  const size_t off_bytes = sizeof(int64_t) * (size_t)(n + 1), aln_bytes = sizeof(Aln) * (size_t)na,
               cig_bytes = 4 * (size_t)ctx->total_cig;
  const size_t at_cig = (aln_bytes + 255) & ~(size_t)255;
  if (ctx->out_buf.size() < at_cig + cig_bytes) ctx->out_buf.resize(at_cig + cig_bytes + (at_cig + cig_bytes) / 4);
  uint8_t *hb = ctx->out_buf.data();
  if (n) {
    if (int rc = stage_d2h(ctx, out->aln_off, ctx->w->aln_off.p, off_bytes)) return rc;
  } else out->aln_off[0] = 0;
  if (na)
    if (int rc = stage_d2h(ctx, hb, ctx->w->alns.p, aln_bytes)) return rc;
  if (cig_bytes)
    if (int rc = stage_d2h(ctx, hb + at_cig, ctx->w->cig_out.p, cig_bytes)) return rc;
  const Aln *alns = (const Aln *)hb;
  const uint32_t *cig = (const uint32_t *)(hb + at_cig);
  const size_t m = (size_t)(na ? na : 1);
  out->rid = (int32_t *)std::malloc(4 * m);
  out->pos = (int64_t *)std::malloc(8 * m);
  out->flag = (int32_t *)std::malloc(4 * m);
  out->is_rev = (int32_t *)std::malloc(4 * m);
  out->score = (int32_t *)std::malloc(4 * m);
  out->nw_score = (int32_t *)std::malloc(4 * m);
  out->nm = (int32_t *)std::malloc(4 * m);
  out->qb = (int32_t *)std::malloc(4 * m);
  out->qe = (int32_t *)std::malloc(4 * m);
  out->rb = (int64_t *)std::malloc(8 * m);
  out->re = (int64_t *)std::malloc(8 * m);
  out->n_cigar = (int32_t *)std::malloc(4 * m);
  out->cigar_off = (int64_t *)std::malloc(8 * m);
  out->n_core = (int32_t *)std::malloc(4 * m);
  out->core_off = (int64_t *)std::malloc(8 * m);
  int64_t words = 0;
  for (int64_t i = 0; i < na; ++i) words += alns[i].n_cigar + alns[i].n_core;
  out->cigar = (uint32_t *)std::malloc(4 * (size_t)(words ? words : 1));
  int64_t w = 0;
  for (int64_t i = 0; i < na; ++i) {
    const Aln &a = alns[i];
    out->rid[i] = a.rid; out->pos[i] = a.pos; out->flag[i] = a.flag; out->is_rev[i] = a.is_rev;
    out->score[i] = a.score; out->nw_score[i] = a.nw_score; out->nm[i] = a.nm;
    out->qb[i] = a.qb; out->qe[i] = a.qe; out->rb[i] = a.rb; out->re[i] = a.re;
    out->n_cigar[i] = a.n_cigar; out->cigar_off[i] = w;
    std::memcpy(out->cigar + w, cig + a.cigar_off, 4 * (size_t)a.n_cigar);
    w += a.n_cigar;
    out->n_core[i] = a.n_core; out->core_off[i] = w;
    std::memcpy(out->cigar + w, cig + a.core_off, 4 * (size_t)a.n_core);
    w += a.n_core;
  }
  out->n_cigar_words = w;
  // MAPQ and MD strings, computed on the device
  out->mapq = (int32_t *)std::malloc(4 * m);
  out->md_off = (int64_t *)std::malloc(8 * (m + 1));
  out->md_off[0] = 0;
  out->n_md_bytes = 0;
  if (na) {
    if (ensure(ctx, ctx->mapq, (size_t)na * 4) || ensure(ctx, ctx->md_len, (size_t)na * 8) ||
        ensure(ctx, ctx->md_off, (size_t)(na + 1) * 8) || ensure(ctx, ctx->fmt_scan_tmp, scan_tmp_bytes(na + 1)))
      return G3A_E_HIP;
    const TextArgs t = text_args(ctx, 1);
    launch_aln_mapq_md(t, na, P<int32_t>(ctx->w->aln_read), P<int32_t>(ctx->mapq), P<int64_t>(ctx->md_len),
                       ctx->stream);
    scan_i64(P<int64_t>(ctx->md_len), P<int64_t>(ctx->md_off), na, ctx->fmt_scan_tmp.p, ctx->fmt_scan_tmp.bytes,
             ctx->stream);
    if (readback(ctx, {P<int64_t>(ctx->md_off) + na})) return G3A_E_HIP;
    const int64_t mdb = ctx->pinned[0];
    if (ensure(ctx, ctx->md, (size_t)mdb + 16)) return G3A_E_HIP;
    launch_aln_md_write(t, na, P<int32_t>(ctx->w->aln_read), P<int64_t>(ctx->md_off), P<char>(ctx->md), ctx->stream);
    HIPCK(hipStreamSynchronize(ctx->stream));
    HIPCK(hipGetLastError());
    out->md = (char *)std::malloc((size_t)mdb + 1);
    if (int rc = stage_d2h(ctx, out->mapq, ctx->mapq.p, (size_t)na * 4)) return rc;
    if (int rc = stage_d2h(ctx, out->md_off, ctx->md_off.p, (size_t)(na + 1) * 8)) return rc;
    if (int rc = stage_d2h(ctx, out->md, ctx->md.p, (size_t)mdb)) return rc;
    out->n_md_bytes = mdb;
  } else {
    out->md = (char *)std::malloc(1);
  }
  out->ms_seed = ctx->ms[0]; out->ms_sa = ctx->ms[1]; out->ms_chain = ctx->ms[2];
  out->ms_extend = ctx->ms[3]; out->ms_traceback = ctx->ms[4]; out->ms_total = ctx->ms[5];
  int64_t c[4] = {0, 0, 0, 0};
  HIPCK(hipMemcpy(c, ctx->w->counters.p, sizeof c, hipMemcpyDeviceToHost));
  int64_t cells = 0;
  HIPCK(hipMemcpy(&cells, P<int64_t>(ctx->w->counters) + 24, 8, hipMemcpyDeviceToHost));
  out->fm_steps = c[0];
  out->ext_cells = cells;
  out->n_seeds = ctx->total_seeds;
  out->n_tasks = ctx->total_tasks;
  out->n_slow_reads = ctx->n_slow;
  return G3A_OK;
}

int g3a_align_batch(g3a_ctx *ctx, int n_reads, const uint8_t *seq, const int32_t *offsets, g3a_batch_out *out) {
  int rc = g3a_memcpy_input(ctx, n_reads, seq, offsets);
  if (rc) return rc;
  rc = g3a_bwa_align(ctx);
  if (rc) return rc;
  return g3a_memcpy_output(ctx, out);
}

int g3a_kernel_ms(const g3a_ctx *ctx, float ms[4]) {
  if (!ctx || !ms) return G3A_E_ARG;
  if (!ctx->has_output) return G3A_E_STATE;
  for (int i = 0; i < 4; ++i) ms[i] = ctx->kms[i];
  return G3A_OK;
}

int g3a_autotune(g3a_ctx *ctx, int *strip_rows, int *lane_blocks_per_cu, float *best_ms) {
  if (!ctx) return G3A_E_ARG;
  if (!ctx->has_output || ctx->w != &ctx->own)
    return fail(ctx, G3A_E_STATE, "autotune needs a batch aligned on the device by g3a_bwa_align");
  HIPCK(hipSetDevice(ctx->device));
  const int form = ctx->form, sr0 = ctx->strip_rows, lb0 = ctx->lane_blocks_per_cu, pk0 = ctx->ext_packed;
  ctx->form = G3A_FORM_BATCHED;
  int64_t *cells = P<int64_t>(ctx->w->counters) + 24;
  // the last batch's tasks: their capacity and their count on the device
  const int64_t cap = ctx->caps.tasks, *n_dev = P<int64_t>(ctx->w->task_off) + ctx->n_reads;
  int64_t saved = 0;
  HIPCK(hipMemcpy(&saved, cells, 8, hipMemcpyDeviceToHost));
  const int srs[2] = {16, 32}, lbs[4] = {2, 4, 8, 16};
  float best = -1;
  int bsr = sr0, blb = lb0, bpk = pk0;
  for (int pk = 0; pk < 2; ++pk)
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 4; ++b) {
      ctx->ext_packed = pk;
      ctx->strip_rows = srs[a];
      ctx->lane_blocks_per_cu = lbs[b];
      float t = 0;
      // first run warms the buffers of this shape, the second is timed
      for (int rep = 0; rep < 2; ++rep) {
        HIPCK(hipEventRecord(ctx->ev[6], ctx->stream));
        if (alloc_ext(ctx, cap)) return G3A_E_HIP;
        if (int rc = run_extension(ctx, cap, n_dev, cells, false)) return rc;
        HIPCK(hipEventRecord(ctx->ev[7], ctx->stream));
        HIPCK(hipEventSynchronize(ctx->ev[7]));
        HIPCK(hipEventElapsedTime(&t, ctx->ev[6], ctx->ev[7]));
      }
      if (best < 0 || t < best) {
        best = t;
        bsr = srs[a];
        blb = lbs[b];
        bpk = pk;
      }
    }
  ctx->strip_rows = bsr;
  ctx->lane_blocks_per_cu = blb;
  ctx->ext_packed = bpk;
  // leave the regions and counters as the aligned batch had them
  if (alloc_ext(ctx, cap)) return G3A_E_HIP;
  if (int rc = run_extension(ctx, cap, n_dev, cells, false)) return rc;
  HIPCK(hipMemcpyAsync(cells, &saved, 8, hipMemcpyHostToDevice, ctx->stream));
  HIPCK(hipStreamSynchronize(ctx->stream));
  ctx->form = form;
  if (strip_rows) *strip_rows = bsr;
  if (lane_blocks_per_cu) *lane_blocks_per_cu = blb;
  if (best_ms) *best_ms = best;
  return G3A_OK;
}

int g3a_set_seed_blocks(g3a_ctx *ctx, int blocks_per_cu) {
  if (!ctx || blocks_per_cu < 0 || blocks_per_cu > 64) return G3A_E_ARG;
  ctx->seed_blocks_per_cu = blocks_per_cu;
  return G3A_OK;
}

int g3a_set_seed_grid(g3a_ctx *ctx, int threads_per_block, int blocks_per_cu) {
  if (!ctx || blocks_per_cu < 0 || blocks_per_cu > 64 || threads_per_block < 64 || threads_per_block > 1024 ||
      threads_per_block % 64)
    return G3A_E_ARG;
  if (threads_per_block != ctx->seed_block) {
    ctx->seed_block = threads_per_block;
    ctx->seed_threads = 0;  // the resident maximum (and the list pool) follow the block size
  }
  ctx->seed_blocks_per_cu = blocks_per_cu;
  return G3A_OK;
}

int g3a_set_chain_serial(g3a_ctx *ctx, int on) {
  if (!ctx || (on != 0 && on != 1)) return G3A_E_ARG;
  ctx->chain_serial = on;
  return G3A_OK;
}

int g3a_set_seed_order(g3a_ctx *ctx, int ordered) {
  if (!ctx || (ordered != 0 && ordered != 1)) return G3A_E_ARG;
  ctx->seed_ordered = ordered;
  return G3A_OK;
}

int g3a_set_learned(g3a_ctx *ctx, int on, g3a_learned_info *info) {
  if (!ctx || on < 0 || on > 2) return G3A_E_ARG;
  if (!ctx->index_on_device) return fail(ctx, G3A_E_STATE, "index not on device");
  DevIndex &x = ctx->idx;
  if (!on) {
    x.saf = nullptr;
    x.keys = nullptr;
    x.li_root = nullptr;
    x.li_leaf = nullptr;
    x.li_knot = nullptr;
    x.li_short = nullptr;
    if (info) *info = ctx->li_info;
    return G3A_OK;
  }
  if (!x.occn) return fail(ctx, G3A_E_ARG, "the learned index needs a text of fewer than 2^32 rows");
  if (!ctx->li_built && ctx->sibling)
    return fail(ctx, G3A_E_STATE, "a sibling context shares its learned index: build it on the source first");
  if (!ctx->li_built) {
    HIPCK(hipSetDevice(ctx->device));
    const int64_t N = x.n + 1;
    if (ensure(ctx, ctx->li_saf, (size_t)N * 4) || ensure(ctx, ctx->li_keys, (size_t)(N + 16) * 8) ||
        ensure(ctx, ctx->li_root, ((size_t)1 << kLiRootBits) * 4 + 4) ||
        ensure(ctx, ctx->li_leaf, sizeof(LiLeaf) * kLiLeaves) || ensure(ctx, ctx->li_leaf_a, (kLiLeaves + 1) * 4) ||
        ensure(ctx, ctx->li_short, 2 * kLiShort * 4) || ensure(ctx, ctx->li_ctr, 64))
      return G3A_E_HIP;
    LearnedBufs b{};
    b.saf = P<uint32_t>(ctx->li_saf);
    b.keys = P<uint64_t>(ctx->li_keys);
    b.root = P<uint32_t>(ctx->li_root);
    b.leaf = P<LiLeaf>(ctx->li_leaf);
    b.leaf_a = P<uint32_t>(ctx->li_leaf_a);
    b.shortt = P<uint32_t>(ctx->li_short);
    b.ctr = P<unsigned long long>(ctx->li_ctr);
    b.n_samples = (int64_t)(ctx->h.sa.size() / ctx->h.sa_width);
    LearnedBuildStats st{};
    HIPCK(hipEventRecord(ctx->ev[6], ctx->stream));
    int64_t n_knots = 0;
    int rc = learned_build_begin(x, b, ctx->stream, &n_knots);
    if (rc == 0) {
      if (ensure(ctx, ctx->li_knot, (size_t)(n_knots + 1) * sizeof(LiKnot))) return G3A_E_HIP;
      b.knot = P<LiKnot>(ctx->li_knot);
      rc = learned_build_end(x, b, ctx->stream, st);
    }
    if (rc == 1) return fail(ctx, G3A_E_ARG, "the learned index needs a text of fewer than 2^32 rows");
    if (rc == 2) return fail(ctx, G3A_E_HIP, "learned index build: HIP error");
    if (rc) {
      char m[256];
      std::snprintf(m, sizeof m,
                    "learned index build: check failed (rows not covered %lld, keys out of order %lld, "
                    "bad walks %lld, short rows %lld)",
                    (long long)st.unfilled_rows, (long long)st.order_violations, (long long)st.bad_walks,
                    (long long)st.short_rows);
      return fail(ctx, G3A_E_INTERNAL, m);
    }
    HIPCK(hipEventRecord(ctx->ev[7], ctx->stream));
    HIPCK(hipEventSynchronize(ctx->ev[7]));
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, ctx->ev[6], ctx->ev[7]));
    g3a_learned_info &li = ctx->li_info;
    li.rows = N;
    li.distinct_keys = st.distinct_keys;
    li.leaves = kLiLeaves;
    li.hidden_units = st.max_knots > 0 ? st.max_knots - 1 : 0;
    li.window_cap = kLiWindowCap;
    li.knots = st.knots;
    li.model_bytes = (int64_t)sizeof(LiLeaf) * kLiLeaves + (int64_t)sizeof(LiKnot) * st.knots +
                     (((int64_t)1 << kLiRootBits) + 1) * 4 + 2 * kLiShort * 4;
    li.sa_bytes = N * 4;
    li.key_bytes = N * 8;
    li.max_err = st.max_err;
    li.keys_over_window = st.keys_over_window;
    li.leaves_over_window = st.leaves_over_window;
    li.mean_err = st.distinct_keys ? (double)st.err_sum / (double)st.distinct_keys : 0.0;
    li.build_ms = ms;
    ctx->li_leaf_w = st.leaf_w;
    ctx->li_built = true;
  }
  x.saf = P<uint32_t>(ctx->li_saf);
  x.keys = P<uint64_t>(ctx->li_keys);
  x.li_root = P<uint32_t>(ctx->li_root);
  x.li_leaf = on == 1 ? P<LiLeaf>(ctx->li_leaf) : nullptr;  // 2: no model lookups
  x.li_knot = P<LiKnot>(ctx->li_knot);
  x.li_short = P<uint32_t>(ctx->li_short);
  x.li_leaf_w = ctx->li_leaf_w;
  if (info) *info = ctx->li_info;
  return G3A_OK;
}

int g3a_learned_stats(const g3a_ctx *ctx, int64_t out[6]) {
  if (!ctx || !out) return G3A_E_ARG;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipMemcpy(out, (const int64_t *)ctx->w->counters.p + 32, 6 * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return G3A_E_HIP;
  return G3A_OK;
}

int g3a_last_intervals(const g3a_ctx *ctx, int32_t *counts, int64_t *kmsn, int64_t cap, int64_t *total) {
  if (!ctx || !counts || !total) return G3A_E_ARG;
  if (!ctx->has_output) return G3A_E_STATE;
  const int n = ctx->n_reads;
  if (n == 0) {
    *total = 0;
    return G3A_OK;
  }
  std::vector<int64_t> base((size_t)n);
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipMemcpy(counts, ctx->w->n_intv.p, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(base.data(), ctx->w->intv_base.p, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return G3A_E_HIP;
  int64_t t = 0, extent = 0;
  for (int r = 0; r < n; ++r) {
    t += counts[r];
    extent = base[r] + counts[r] > extent ? base[r] + counts[r] : extent;
  }
  *total = t;
  if (!kmsn) return G3A_OK;
  if (cap < t) return G3A_E_ARG;
  std::vector<Intv> iv((size_t)extent);
  if (extent && hipMemcpy(iv.data(), ctx->w->intv.p, (size_t)extent * sizeof(Intv), hipMemcpyDeviceToHost) != hipSuccess)
    return G3A_E_HIP;
  int64_t w = 0;
  for (int r = 0; r < n; ++r)
    for (int k = 0; k < counts[r]; ++k) {
      const Intv &v = iv[(size_t)(base[r] + k)];
      kmsn[4 * w] = v.k;
      kmsn[4 * w + 1] = v.s;
      kmsn[4 * w + 2] = v.m;
      kmsn[4 * w + 3] = v.n;
      ++w;
    }
  return G3A_OK;
}

int g3a_autotune_seed(g3a_ctx *ctx, int *threads_per_block, int *blocks_per_cu, int *ordered, float *best_ms) {
  if (!ctx) return G3A_E_ARG;
  if (!ctx->has_output || ctx->w != &ctx->own)
    return fail(ctx, G3A_E_STATE, "autotune needs a batch aligned on the device by g3a_bwa_align");
  // grids of 16 waves per compute unit, from one block (one LDS copy of the
  // superblock counts) to six smaller ones
  const int grids[4][2] = {{1024, 1}, {512, 2}, {256, 4}, {128, 6}};
  float best = -1;
  int bt = ctx->seed_block, bb = ctx->seed_blocks_per_cu, bo = ctx->seed_ordered;
  const int graph = ctx->use_graph;
  ctx->use_graph = 0;  // the stage times come from the eager launches' events
  for (int o = 0; o < 2; ++o)
    for (const auto &g : grids) {
      g3a_set_seed_grid(ctx, g[0], g[1]);
      ctx->seed_ordered = o;
      if (int rc = g3a_bwa_align(ctx)) {
        ctx->use_graph = graph;
        return rc;
      }
      // the seeding stage: the order's key and sort kernels and k_seed
      if (best < 0 || ctx->ms[0] < best) {
        best = ctx->ms[0];
        bt = g[0];
        bb = g[1];
        bo = o;
      }
    }
  g3a_set_seed_grid(ctx, bt, bb);
  ctx->seed_ordered = bo;
  ctx->use_graph = graph;
  if (threads_per_block) *threads_per_block = bt;
  if (blocks_per_cu) *blocks_per_cu = bb;
  if (ordered) *ordered = bo;
  if (best_ms) *best_ms = best;
  return G3A_OK;
}

int g3a_set_graph(g3a_ctx *ctx, int on) {
  if (!ctx || (on != 0 && on != 1)) return G3A_E_ARG;
  ctx->use_graph = on;
  if (!on) {
    HIPCK(hipSetDevice(ctx->device));
    free_graph(ctx);
  }
  return G3A_OK;
}

int g3a_set_batch_caps(g3a_ctx *ctx, const double v[7]) {
  if (!ctx || !v) return G3A_E_ARG;
  for (int i = 0; i < 7; ++i)
    if (!(v[i] >= 0) || v[i] > 1e9) return G3A_E_ARG;
  for (int i = 0; i < 4; ++i) ctx->per_read[i] = v[i];
  ctx->per_base[0] = v[4];
  ctx->per_base[1] = v[5];
  ctx->slow_cap = (int)v[6];
  return G3A_OK;
}

int g3a_graph_info(const g3a_ctx *ctx, int64_t info[3]) {
  if (!ctx || !info) return G3A_E_ARG;
  info[0] = ctx->graph_used ? 1 : 0;
  info[1] = ctx->graph_captures;
  info[2] = ctx->overflow_reruns;
  return G3A_OK;
}

int g3a_set_form(g3a_ctx *ctx, int form) {
  if (!ctx || (form != G3A_FORM_PLAIN && form != G3A_FORM_BATCHED)) return G3A_E_ARG;
  ctx->form = form;
  return G3A_OK;
}

int g3a_set_tuning(g3a_ctx *ctx, int strip_rows, int lane_blocks_per_cu, int wave_blocks_per_cu) {
  if (!ctx || (strip_rows != 16 && strip_rows != 32) || lane_blocks_per_cu < 1 || lane_blocks_per_cu > 64 ||
      wave_blocks_per_cu < 1 || wave_blocks_per_cu > 64)
    return G3A_E_ARG;
  ctx->strip_rows = strip_rows;
  ctx->lane_blocks_per_cu = lane_blocks_per_cu;
  ctx->wave_blocks_per_cu = wave_blocks_per_cu;
  return G3A_OK;
}

int g3a_set_wide_layout(g3a_ctx *ctx, int on) {
  if (!ctx || (on != 0 && on != 1)) return G3A_E_ARG;
  if (ctx->index_on_device) return fail(ctx, G3A_E_STATE, "the layout is chosen before g3a_memcpy_index");
  ctx->force_wide = on;
  return G3A_OK;
}

int g3a_set_ext_packed(g3a_ctx *ctx, int packed) {
  if (!ctx || (packed != 0 && packed != 1)) return G3A_E_ARG;
  ctx->ext_packed = packed;
  return G3A_OK;
}

int g3a_ext_packed(const g3a_ctx *ctx) { return ctx ? ctx->ext_packed : -1; }

int g3a_last_stats(const g3a_ctx *ctx, float ms[6], int64_t work[6]) {
  if (!ctx || !ms || !work) return G3A_E_ARG;
  if (!ctx->has_output) return G3A_E_STATE;
  for (int i = 0; i < 6; ++i) ms[i] = ctx->ms[i];
  int64_t c[32];
  if (hipMemcpy(c, ctx->w->counters.p, sizeof c, hipMemcpyDeviceToHost) != hipSuccess) return G3A_E_HIP;
  work[0] = ctx->total_seeds;
  work[1] = ctx->total_tasks;
  work[2] = c[24];
  work[3] = c[0];
  work[4] = ctx->n_slow;
  work[5] = ctx->total_alns;
  return G3A_OK;
}

void g3a_batch_out_free(g3a_batch_out *o) {
  if (!o) return;
  std::free(o->aln_off); std::free(o->rid); std::free(o->pos); std::free(o->flag); std::free(o->is_rev);
  std::free(o->score); std::free(o->nw_score); std::free(o->nm); std::free(o->qb); std::free(o->qe);
  std::free(o->rb); std::free(o->re); std::free(o->n_cigar); std::free(o->cigar_off); std::free(o->n_core);
  std::free(o->core_off); std::free(o->cigar); std::free(o->mapq); std::free(o->md_off); std::free(o->md);
  std::memset(o, 0, sizeof *o);
}

void g3a_destroy(g3a_ctx *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  free_graph(ctx);
  DevBuf *bufs[] = {&ctx->occ, &ctx->kmer, &ctx->occn, &ctx->sbc, &ctx->pac, &ctx->sa, &ctx->ctg_off, &ctx->ctg_len, &ctx->ctg_alt, &ctx->seq, &ctx->off,
                    &ctx->own.intv, &ctx->own.n_intv, &ctx->own.n_seed, &ctx->own.slow, &ctx->own.intv_base, &ctx->own.seed_off,
                    &ctx->own.counters, &ctx->own.slow_list, &ctx->own.slow_scratch, &ctx->own.slow_n, &ctx->own.slow_off, &ctx->own.seeds,
                    &ctx->own.seeds2, &ctx->own.ch, &ctx->own.nxt, &ctx->own.order, &ctx->own.tmpi, &ctx->own.w16, &ctx->own.kept, &ctx->own.nodes,
                    &ctx->own.tasks_tmp, &ctx->own.n_task, &ctx->own.task_off, &ctx->own.chain_big, &ctx->own.tasks, &ctx->own.regs, &ctx->own.flag, &ctx->own.sec,
                    &ctx->own.n_aln, &ctx->own.aln_off, &ctx->own.aln_reg, &ctx->own.aln_read, &ctx->own.aln_rank, &ctx->own.z_size,
                    &ctx->own.cig_size, &ctx->own.z_off, &ctx->own.cig_off, &ctx->own.z, &ctx->own.cigar, &ctx->own.alns, &ctx->own.scan_tmp,
                    &ctx->own.cig_words, &ctx->own.cig_coff, &ctx->own.cig_out, &ctx->own.seed_key, &ctx->own.seed_key2, &ctx->own.seed_idx,
                    &ctx->own.seed_order, &ctx->own.seed_sort_tmp,
                    &ctx->own.ext_res[0], &ctx->own.ext_res[1], &ctx->own.ext_key, &ctx->own.ext_hist, &ctx->own.totals,
                    &ctx->own.ext_sorted, &ctx->own.ext_slow, &ctx->own.ext_ctr, &ctx->own.ext_rowbuf,
                    &ctx->own.lep_pool, &ctx->own.eh_pool, &ctx->own.tb_order, &ctx->own.tb_ctr, &ctx->own.tb_nw, &ctx->li_saf, &ctx->li_keys,
                    &ctx->li_root, &ctx->li_leaf, &ctx->li_knot, &ctx->li_leaf_a, &ctx->li_short, &ctx->li_ctr,
                    &ctx->ctg_names, &ctx->ctg_name_off, &ctx->log_tab, &ctx->names, &ctx->name_off, &ctx->quals,
                    &ctx->qual_off, &ctx->txt_len, &ctx->txt_off, &ctx->txt, &ctx->txt_bound, &ctx->stage_off, &ctx->txt_stage, &ctx->fmt_flag, &ctx->mapq, &ctx->md_len,
                    &ctx->md_off, &ctx->md, &ctx->fmt_scan_tmp, &ctx->pe_hist, &ctx->pe_stat, &ctx->pe_v, &ctx->pe_out};
  for (DevBuf *b : bufs) free_buf(*b);
  for (auto &ev : ctx->ev)
    if (ev) hipEventDestroy(ev);
  for (auto &ev : ctx->kev)
    if (ev) hipEventDestroy(ev);
  if (ctx->pinned) hipHostFree(ctx->pinned);
  for (int k = 0; k < 2; ++k) {
    if (ctx->stage[k]) hipHostFree(ctx->stage[k]);
    if (ctx->stage_ev[k]) hipEventDestroy(ctx->stage_ev[k]);
  }
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
}

}  // extern "C"
